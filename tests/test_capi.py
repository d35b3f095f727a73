"""The C-ABI library loads and exports every function the public headers
declare (no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

import turbobfs_amd
from tests.conftest import ROOT

_DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def _declared_functions():
    names = {}
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//.*", "", text)
        text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
        for m in _DECL.finditer(text):
            name = m.group(1)
            if name in ("if", "while", "for", "return", "sizeof"):
                continue
            names[name] = os.path.basename(h)
    return names


def test_headers_declare_reference_entry_points():
    names = _declared_functions()
    for f in ("bfs_gpu_td_csc_sc", "bfs_gpu_td_csc_wa", "bfs_gpu_bu_csc_sc",
              "bfs_gpu_bu_csc_wa", "bfs_gpu_tdbu_csc_sc", "bfs_gpu_tdbu_csc_wa",
              "bfs_gpu_td_cooc_sc"):
        assert f in names, f


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(turbobfs_amd.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_every_symbol():
    from turbobfs_amd import _lib

    missing = [n for n in _declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing
    turbobfs_amd.load()


def test_status_strings_and_defaults():
    lib = turbobfs_amd.load()
    assert lib.tbfs_status_str(0) == b"ok"
    assert lib.tbfs_status_str(5).startswith(b"path count")
    o = turbobfs_amd.default_opts()
    assert o.mode == 0 and o.sigma == 1
    # the tunables are left at 0: the autotuner's value, else the default
    assert o.alpha == 0 and o.beta == 0 and o.stage_ratio == 0 and o.lb_grid == 0


def test_python_structs_match_the_header(tmp_path):
    """Every struct the binding mirrors has the header's size and field
    offsets (a C program built from include/tbfs.h prints them)."""
    import subprocess

    from turbobfs_amd import _lib

    pairs = [("tbfs_opts", _lib.Opts), ("tbfs_stats", _lib.Stats),
             ("tbfs_server_opts", _lib.ServerOpts), ("tbfs_query_done", _lib.QueryDone),
             ("tbfs_degree_stats", _lib.DegreeStats)]
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "tbfs.h"', "int main(void) {"]
    for cname, py in pairs:
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe),
                    str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(x.split()[:2]): int(x.split()[2]) for x in out if x.strip()}
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
