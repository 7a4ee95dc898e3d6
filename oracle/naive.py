"""ctypes wrapper of naive.c (the reference's loop nest) -- test infrastructure only."""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path
from typing import Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "_build" / "libnaive.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE), "_build/libnaive.so"], check=True)
    return _LIB


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB))
    return _lib


def _strides(labels: Sequence[str], extents: Sequence[int]):
    st, s = {}, 1
    for l, e in zip(labels, extents):
        st[l] = s
        s *= e
    return st


def naive_contract(c, a, b, sizes, A, B, C, alpha=1.0, beta=0.0):
    """C (TCCG layout numpy array, float32 or float64) updated in place by the
    reference's loop nest; c/a/b are index label lists, sizes a dict."""
    lib = _load()
    dt = C.dtype
    if dt not in (np.float32, np.float64):
        raise TypeError(dt)
    sa = _strides(a, [sizes[l] for l in a])
    sb = _strides(b, [sizes[l] for l in b])
    sc = _strides(c, [sizes[l] for l in c])
    free = list(c)
    kset = [l for l in a if l in b and l not in c]
    i64 = lambda v: (ctypes.c_int64 * max(1, len(v)))(*v)
    fext = i64([sizes[l] for l in free])
    fsa = i64([sa.get(l, 0) for l in free])
    fsb = i64([sb.get(l, 0) for l in free])
    fsc = i64([sc[l] for l in free])
    kext = i64([sizes[l] for l in kset])
    ksa = i64([sa[l] for l in kset])
    ksb = i64([sb[l] for l in kset])
    A = np.ascontiguousarray(A, dtype=dt)
    B = np.ascontiguousarray(B, dtype=dt)
    assert C.flags.c_contiguous
    ptr = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    if dt == np.float32:
        fn, T = lib.naive_contract_f32, ctypes.c_float
    else:
        fn, T = lib.naive_contract_f64, ctypes.c_double
    fn.argtypes = [ctypes.c_void_p] * 3 + [T, T, ctypes.c_int] + [ctypes.c_void_p] * 4 + [ctypes.c_int] + [ctypes.c_void_p] * 3
    fn(ptr(A), ptr(B), ptr(C), alpha, beta, len(free), fext, fsa, fsb, fsc, len(kset), kext, ksa, ksb)
    return C
