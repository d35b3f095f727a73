"""The project's argv-compatible drivers (turbobfs_amd/tbfs_exec_{td,bu,tdbu,
cooc}, built from turbobfs_amd/csrc/tbfs_exec.c without the reference tree):
the reference's run-script argument sets (TurboBFS_CSC_TD/run_bfstdcsc.sh
and the other variants' scripts), its "... are correct" check lines, and the
printed S / sigma tables (printBFS format, TurboBFS_CSC_TD/utils_bfs.c:26-90)
against the oracle; and the argument count checked before any argument is
read (the reference's TD main reads argv[8] with argc checked against 8,
TurboBFS_CSC_TD/bfstdcsc_main.c:82-97)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.conftest import ROOT, golden
from tests.test_reference_drivers import CASES, S_OK, SIGMA_OK, _matrix
from turbobfs_amd import mtx

BIN = os.path.join(ROOT, "turbobfs_amd")


def _exe(variant):
    path = os.path.join(BIN, "tbfs_exec_" + variant)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (make -C turbobfs_amd/csrc)")
    return path


@pytest.mark.parametrize("variant,nargs", [("td", 8), ("bu", 9), ("tdbu", 9), ("cooc", 7)])
def test_argument_count_checked_first(variant, nargs):
    """One argument short: a message and exit status 1, no crash."""
    args = ["x"] * (nargs - 1)
    proc = subprocess.run([_exe(variant), *args], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 1
    assert f"should be {nargs}" in proc.stdout
    assert "Usage" in proc.stderr


def _parse_tables(out: str):
    """Rows of the printed "S   S_hgpu" and sigma tables: {i: (a, b)}."""
    S, sig = {}, {}
    cur = None
    for line in out.splitlines():
        if line.startswith("S   S_hgpu"):
            cur = S
            continue
        if line.startswith("sigma"):
            cur = sig
            continue
        if cur is None or not line.strip():
            if not line.strip():
                cur = None
            continue
        parts = line.split(",")
        if len(parts) >= 2 and parts[0].lstrip("-").isdigit():
            cur[int(parts[0])] = tuple(float(x) for x in parts[1:])
    return S, sig


OURS = [(c[0].replace("exec_", ""), c[1], c[2], c[3]) for c in CASES]


@pytest.mark.gpu
@pytest.mark.parametrize("case", OURS, ids=[f"{c[0]}-{c[1]}-" + c[2].replace(" ", "_")
                                            for c in OURS])
def test_run_script_argument_sets(gpu, case, tmp_path_factory):
    variant, name, argv, expect = case
    path = _matrix(name, tmp_path_factory.mktemp("mtx"))
    args = argv.split()
    print(f"[exec] {variant} {name} {argv}", flush=True)
    proc = subprocess.run([_exe(variant), path, *args], capture_output=True, text=True,
                          timeout=300)
    assert proc.returncode == 0, proc.stderr[-2000:]
    for line in expect:
        assert line in proc.stdout, proc.stdout[-3000:]
    assert "error in" not in proc.stdout, proc.stdout[-3000:]
    # the printed tables against the oracle (printing on: p = 1); argv after
    # the file: w ug [format] p r ...
    p_idx = 2 if variant == "cooc" else 3
    if args[p_idx] != "1":
        return
    ug = bool(int(args[1]))
    r = int(args[p_idx + 1])
    g = mtx.read_csc(golden(name + ".mtx") if os.path.exists(golden(name + ".mtx"))
                     else golden(name + ".mtx.gz"), ug=ug)
    S, sig, ovf, _ = oracle.textbook(g.CP, g.IC, r)
    S_tab, sig_tab = _parse_tables(proc.stdout)
    if variant != "cooc":
        assert S_tab, proc.stdout[-2000:]
        for i, (s_seq, s_gpu) in S_tab.items():
            want = (0 if S[i] < 0 else S[i]) if variant == "td" else S[i]
            assert s_gpu == want and s_seq == want, (i, s_seq, s_gpu, want)
    if variant in ("td", "cooc"):
        assert sig_tab, proc.stdout[-2000:]
        for i, (a, b) in sig_tab.items():
            assert a == b == float(np.float32(sig[i])), (i, a, b, sig[i])
