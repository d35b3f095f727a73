"""The kernels declared by include/bfsgputdcsc.cuh and include/bfsgputdcooc.cuh
(the reference's .cuh prototypes, TurboBFS_CSC_TD/bfsgputdcsc.cuh:18-19 and
TurboBFS_COOC_TD/bfsgputdcooc.cuh:18-19), launched by a separate HIP program
(turbobfs_amd/cuh_check, built with the library) through those headers."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "turbobfs_amd", "cuh_check")
FT = np.array([0, 2, 0, 1, 5, 0, 3, 0])
SIGMA0 = np.array([1, 0, 0, 2, 0, 1, 0, 0], dtype=float)


def _parse(line):
    name, rest = line.split(" ", 1)
    out = {"name": name}
    for item in rest.split():
        k, v = item.split("=")
        out[k] = np.array([float(x) for x in v.split(",")]) if "," in v else float(v)
    return out


def test_cuh_kernels_through_headers(gpu):
    assert os.path.exists(EXE), "cuh_check is built by make -C turbobfs_amd/csrc"
    proc = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 0, proc.stderr
    csc, cooc = [_parse(x) for x in proc.stdout.strip().splitlines()]
    # CSC (bfsgputdcsc_sc.cu:194-208): f = ft where ft > 0 (else 0); those get
    # sigma += ft, S = d (7 here), and c = 1
    hit = FT > 0
    assert csc["name"] == "csc" and csc["c"] == 1
    np.testing.assert_array_equal(csc["f"], np.where(hit, FT, 0))
    np.testing.assert_array_equal(csc["sigma"], SIGMA0 + np.where(hit, FT, 0))
    np.testing.assert_array_equal(csc["S"], np.where(hit, 7, 0))
    # COOC (bfsgputdcooc_sc.cu:175-188): f = ft only where sigma is still 0
    fo = np.where(SIGMA0 < 0.01, FT, 0)
    assert cooc["name"] == "cooc" and cooc["c"] == 1
    np.testing.assert_array_equal(cooc["f"], fo)
    np.testing.assert_array_equal(cooc["sigma"], SIGMA0 + np.where(fo > 0, fo, 0))
    np.testing.assert_array_equal(cooc["S"], np.zeros(8))
