"""Write the golden fixtures under tests/golden/ (run in the build container).

Inputs come from this project's deterministic generators; expected outputs
come from the reference's own sequential C code (bfs_seq_td_csc and
bfs_seq_td_cooc, compiled from the reference sources into oracle/_ref by
`make -C oracle ref`).  The reference is not needed afterwards: the fixtures
are plain arrays, so the GPU tests can use them where the reference tree is
absent.  Each fixture also stores the SHA-256 of the generated CSC so a change
in the generators is caught.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from turbobfs_amd import graphs, mtx  # noqa: E402


def csc_sha(g) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(g.CP, np.int32).tobytes())
    h.update(np.ascontiguousarray(g.IC, np.int32).tobytes())
    return h.hexdigest()


def ref_outputs(g, roots):
    S_td, sig_td, S_bu, sig_tdbu, sig_cooc = [], [], [], [], []
    I, J = graphs.csc_to_cooc(g)
    for r in roots:
        s, f = oracle.ref_seq_td_csc(g.CP, g.IC, int(r), "td")
        S_td.append(s)
        sig_td.append(f)
        s, _ = oracle.ref_seq_td_csc(g.CP, g.IC, int(r), "bu")
        S_bu.append(s)
        _, si = oracle.ref_seq_td_csc(g.CP, g.IC, int(r), "tdbu")
        sig_tdbu.append(si)
        sig_cooc.append(oracle.ref_seq_td_cooc(I, J, g.n, int(r)))
    return dict(roots=np.asarray(roots, np.int32), S_td=np.array(S_td),
                sigma_td=np.array(sig_td), S_bu=np.array(S_bu),
                sigma_tdbu=np.array(sig_tdbu), sigma_cooc=np.array(sig_cooc))


def main():
    if not oracle.ref_available("td"):
        sys.exit("oracle/_ref is not built: run make -C oracle ref")
    # ER-4096: undirected G(4096, 8/4095), seed 1, source 0; the CSC is stored
    g = graphs.erdos_renyi(4096, 8.0 / 4095.0, seed=1)
    out = ref_outputs(g, [0])
    np.savez_compressed(os.path.join(HERE, "er4096_seed1.npz"), CP=g.CP, IC=g.IC,
                        sha256=np.array(csc_sha(g)), **out)
    # R-MAT scale 12, edge factor 16, seed 2: only the CSC hash is stored
    g = graphs.rmat(12, 16, 0.57, 0.19, 0.19, seed=2, relabel=True)
    roots = graphs.sample_sources(graphs.nonzero_degree_vertices(g), 4, seed=2)
    out = ref_outputs(g, roots)
    np.savez_compressed(os.path.join(HERE, "rmat12_seed2.npz"), sha256=np.array(csc_sha(g)),
                        n=g.n, nz=g.nz, **out)
    # luxembourg_osm (reference data file), source 0, 1035 depths
    g = mtx.read_csc(os.path.join(HERE, "luxembourg_osm.mtx.gz"))
    out = ref_outputs(g, [0])
    np.savez_compressed(os.path.join(HERE, "luxembourg_osm_r0.npz"),
                        sha256=np.array(csc_sha(g)), **out)
    for f in ("er4096_seed1.npz", "rmat12_seed2.npz", "luxembourg_osm_r0.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
