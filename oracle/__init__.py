"""CPU checker for the GPU BFS.  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package; the product library never does.

- liboracle.so: this project's C restatement (oracle.c) of the reference's
  sequential BFS (TurboBFS_CSC_TD/bfstdcsc_seq.c:27-96 and friends), with
  path counts held as (uint64 mod 2^64, overflow flag).
- _ref/*.so: the reference's own sequential C files compiled in place
  (oracle/Makefile `ref` target); present only where the reference tree was
  available at build time.  Float/int path counts, as the reference has them.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_u8p = C.POINTER(C.c_uint8)
_f32p = C.POINTER(C.c_float)

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        lib.orc_bfs_seq_td_csc_passes.restype = C.c_int
        lib.orc_bfs_seq_td_csc_passes.argtypes = [_i32p, _i32p, C.c_int32, C.c_int32, _i32p,
                                                  _u64p, _u8p, C.c_int32]
        lib.orc_bfs_seq_td_csc.restype = C.c_int
        lib.orc_bfs_seq_td_csc.argtypes = [_i32p, _i32p, C.c_int32, C.c_int32, C.c_int32,
                                           _i32p, _u64p, _u8p]
        lib.orc_bfs_seq_td_cooc.restype = C.c_int
        lib.orc_bfs_seq_td_cooc.argtypes = [_i32p, _i32p, C.c_int32, C.c_int32, C.c_int32,
                                            _u64p, _u8p]
        lib.orc_bfs_textbook.restype = C.c_int
        lib.orc_bfs_textbook.argtypes = [_i32p, _i32p, C.c_int32, C.c_int32, _i32p, _u64p, _u8p]
        lib.orc_count_exceed.restype = C.c_int64
        lib.orc_count_exceed.argtypes = [_u64p, _u8p, C.c_int32, C.c_uint64]
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _csc(CP, IC):
    return (np.ascontiguousarray(CP, dtype=np.int32), np.ascontiguousarray(IC, dtype=np.int32))


def seq_td_csc(CP, IC, r, s_unreached=-1):
    """Restated bfs_seq_td_csc.  Returns (S, sigma, ovf, passes)."""
    CP, IC = _csc(CP, IC)
    n = CP.shape[0] - 1
    S = np.empty(n, np.int32)
    sig = np.empty(n, np.uint64)
    ovf = np.empty(n, np.uint8)
    d = load().orc_bfs_seq_td_csc(_p(IC, _i32p), _p(CP, _i32p), n, int(r), int(s_unreached),
                                  _p(S, _i32p), _p(sig, _u64p), _p(ovf, _u8p))
    if d < 0:
        raise ValueError("orc_bfs_seq_td_csc: bad arguments")
    return S, sig, ovf, d


def seq_td_csc_passes(CP, IC, r, max_passes):
    """The restated bfs_seq_td_csc loop stopped after max_passes passes (a
    timing sample of a BFS too long for one CPU thread).  Returns
    (S, sigma, ovf, passes) of the partial run; S uses -1 for unreached."""
    CP, IC = _csc(CP, IC)
    n = CP.shape[0] - 1
    S = np.empty(n, np.int32)
    sig = np.empty(n, np.uint64)
    ovf = np.empty(n, np.uint8)
    d = load().orc_bfs_seq_td_csc_passes(_p(IC, _i32p), _p(CP, _i32p), n, int(r), _p(S, _i32p),
                                         _p(sig, _u64p), _p(ovf, _u8p), int(max_passes))
    if d < 0:
        raise ValueError("orc_bfs_seq_td_csc_passes: bad arguments")
    return S, sig, ovf, d


def seq_td_cooc(I, J, n, r):
    """Restated bfs_seq_td_cooc.  Returns (sigma, ovf, passes)."""
    I = np.ascontiguousarray(I, dtype=np.int32)
    J = np.ascontiguousarray(J, dtype=np.int32)
    sig = np.empty(n, np.uint64)
    ovf = np.empty(n, np.uint8)
    d = load().orc_bfs_seq_td_cooc(_p(I, _i32p), _p(J, _i32p), I.shape[0], n, int(r),
                                   _p(sig, _u64p), _p(ovf, _u8p))
    if d < 0:
        raise ValueError("orc_bfs_seq_td_cooc: bad arguments")
    return sig, ovf, d


def textbook(CP, IC, r):
    """O(n+m) queue BFS with path counting.  Returns (S, sigma, ovf, levels)."""
    CP, IC = _csc(CP, IC)
    n = CP.shape[0] - 1
    S = np.empty(n, np.int32)
    sig = np.empty(n, np.uint64)
    ovf = np.empty(n, np.uint8)
    lv = load().orc_bfs_textbook(_p(CP, _i32p), _p(IC, _i32p), n, int(r), _p(S, _i32p),
                                 _p(sig, _u64p), _p(ovf, _u8p))
    if lv < 0:
        raise ValueError("orc_bfs_textbook: bad arguments")
    return S, sig, ovf, lv


def count_exceed(sigma, ovf, limit):
    sigma = np.ascontiguousarray(sigma, dtype=np.uint64)
    ovf = np.ascontiguousarray(ovf, dtype=np.uint8)
    return int(load().orc_count_exceed(_p(sigma, _u64p), _p(ovf, _u8p), sigma.shape[0],
                                       int(limit)))


# ------------------------------------------------ the reference's own code ---
def ref_available(name: str = "td") -> bool:
    return os.path.exists(os.path.join(REF_DIR, f"libref_{name}.so"))


_ref = {}


def _ref_lib(name):
    if name not in _ref:
        path = os.path.join(REF_DIR, f"libref_{name}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        _ref[name] = C.CDLL(path)
    return _ref[name]


def ref_seq_td_csc(CP, IC, r, variant="td"):
    """The reference's bfs_seq_td_csc.  variant 'td' (float sigma, S left at
    its calloc value 0), 'bu' (float sigma, S=-1/S[r]=0) or 'tdbu' (int
    sigma, S=-1/S[r]=0).  Returns (S, sigma)."""
    CP, IC = _csc(CP, IC)
    n = CP.shape[0] - 1
    lib = _ref_lib(variant)
    S = np.zeros(n, np.int32)
    if variant == "tdbu":
        sig = np.zeros(n, np.int32)
        sp = _p(sig, _i32p)
    else:
        sig = np.zeros(n, np.float32)
        sp = _p(sig, _f32p)
    lib.bfs_seq_td_csc.restype = C.c_int
    lib.bfs_seq_td_csc(_p(IC, _i32p), _p(CP, _i32p), _p(S, _i32p), sp, C.c_int(int(r)),
                       C.c_int(IC.shape[0]), C.c_int(n))
    return S, sig


def ref_seq_td_cooc(I, J, n, r):
    I = np.ascontiguousarray(I, dtype=np.int32)
    J = np.ascontiguousarray(J, dtype=np.int32)
    lib = _ref_lib("cooc")
    sig = np.zeros(n, np.float32)
    lib.bfs_seq_td_cooc(_p(I, _i32p), _p(J, _i32p), _p(sig, _f32p), C.c_int(int(r)),
                        C.c_int(I.shape[0]), C.c_int(n))
    return sig
