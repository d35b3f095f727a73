"""The reference's own command-line drivers, linked against libturbobfs.so.

`make -C oracle dropin` compiles every .c file of a reference variant
directory (its main, sequential BFS, Matrix Market reader, format transforms,
degree statistics and checkers) where it lies and links it against this
library in place of the reference's nvcc-built .cu objects, into
oracle/_ref/exec_{td,bu,tdbu,cooc}.  Each driver runs its own sequential BFS,
calls the GPU entry point exported here, and compares the two with its own
checkers (TurboBFS_CSC_TD/utils_bfs.c:93-145), printing "... are correct".

The argument sets are those of the reference's run scripts
(TurboBFS_CSC_TD/run_bfstdcsc.sh, TurboBFS_CSC_BU/run_bfsbucsc.sh,
TurboBFS_CSC_TDBU/run_bfstdbucsc.sh, TurboBFS_COOC_TD/run_bfstdcooc.sh):
argv = file w ug format p r repet seq [bbreak] (COOC has no format), on the
matrices bundled with the reference (copies under tests/golden/).
"""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT, golden

REF_BIN = os.path.join(ROOT, "oracle", "_ref")

SIGMA_OK = "values of the GPU computed sigma vector are correct"
S_OK = "values of the GPU computed S vector are correct"

# (driver, matrix, argv after the file name, lines the driver must print)
CASES = []
for fmt in (0, 1):
    CASES += [
        ("exec_td", "mycielskian4", f"0 1 {fmt} 1 0 1 1", (SIGMA_OK, S_OK)),
        ("exec_td", "Tina_AskCal", f"0 0 {fmt} 1 10 1 1", (SIGMA_OK, S_OK)),
        ("exec_td", "Ragusa18", f"1 0 {fmt} 1 0 10 1", (SIGMA_OK, S_OK)),
        ("exec_td", "cage4", f"1 0 {fmt} 1 0 10 1", (SIGMA_OK, S_OK)),
        ("exec_td", "luxembourg_osm", f"0 1 {fmt} 0 0 2 1", (SIGMA_OK, S_OK)),
    ]
    for bb in (0, 1):
        CASES += [
            ("exec_bu", "mycielskian4", f"0 1 {fmt} 1 0 1 1 {bb}", (S_OK,)),
            ("exec_bu", "Tina_AskCal", f"0 0 {fmt} 1 10 1 1 {bb}", (S_OK,)),
            ("exec_bu", "Ragusa18", f"1 0 {fmt} 1 0 1 1 {bb}", (S_OK,)),
            ("exec_bu", "cage4", f"1 0 {fmt} 1 0 1 1 {bb}", (S_OK,)),
        ]
    CASES += [
        ("exec_tdbu", "mycielskian4", f"0 1 {fmt} 1 0 1 1 1", (S_OK,)),
        ("exec_tdbu", "Tina_AskCal", f"0 0 {fmt} 1 10 1 1 1", (S_OK,)),
        ("exec_tdbu", "Ragusa18", f"1 0 {fmt} 1 0 10 1 1", (S_OK,)),
        ("exec_tdbu", "cage4", f"1 0 {fmt} 1 0 10 1 1", (S_OK,)),
        ("exec_tdbu", "luxembourg_osm", f"0 1 {fmt} 0 0 2 1 1", (S_OK,)),
    ]
CASES += [
    ("exec_cooc", "mycielskian4", "0 1 1 0 1 1", (SIGMA_OK,)),
    ("exec_cooc", "Tina_AskCal", "0 0 1 10 1 1", (SIGMA_OK,)),
    ("exec_cooc", "Ragusa18", "1 0 1 0 10 1", (SIGMA_OK,)),
    ("exec_cooc", "cage4", "1 0 1 0 10 1", (SIGMA_OK,)),
    ("exec_cooc", "luxembourg_osm", "0 1 0 0 2 1", (SIGMA_OK,)),
]


def _matrix(name: str, tmp_dir) -> str:
    path = golden(name + ".mtx")
    if os.path.exists(path):
        return path
    out = os.path.join(str(tmp_dir), name + ".mtx")
    if not os.path.exists(out):
        with gzip.open(golden(name + ".mtx.gz"), "rb") as src, open(out, "wb") as dst:
            shutil.copyfileobj(src, dst)
    return out


@pytest.fixture(scope="module")
def mtx_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("mtx")


def _ids(case):
    return f"{case[0]}-{case[1]}-" + case[2].replace(" ", "_")


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_reference_driver(gpu, case, mtx_dir):
    exe, name, argv, expect = case
    path = os.path.join(REF_BIN, exe)
    if not os.path.exists(path):
        pytest.skip(f"{exe} not built (make -C oracle dropin needs the reference tree)")
    mtx = _matrix(name, mtx_dir)
    print(f"[driver] {exe} {name} {argv}", flush=True)
    proc = subprocess.run([path, mtx, *argv.split()], capture_output=True,
                          text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr[-2000:]
    assert "turbobfs:" not in proc.stderr, proc.stderr[-2000:]
    for line in expect:
        assert line in proc.stdout, proc.stdout[-3000:]
    assert "error in" not in proc.stdout, proc.stdout[-3000:]


def test_driver_cases_cover_every_entry_point():
    """Both formats of every variant appear, so all seven entry points run."""
    seen = {(c[0], c[2].split()[2]) for c in CASES if c[0] != "exec_cooc"}
    for exe in ("exec_td", "exec_bu", "exec_tdbu"):
        assert (exe, "0") in seen and (exe, "1") in seen
    assert any(c[0] == "exec_cooc" for c in CASES)
