"""ctypes driver of avx2_gett.cpp, the restated AVX2 GETT -- test infrastructure
and CPU baseline only.

Classification (m/n/k sets, A/B swap so that the internal A holds C's
stride-one index) follows tccg/tccg_util.py:212-296, tccg/gett.py:64-89;
block sizes follow the reference's policy for variant 0 (n-outer):
mr=24/nr=4 (float), 12/4 (double), mc from getPossibleValues(mr, M, 192|96, 3),
nc = N when N < 4600, kc from getPossibleValues(1, K, 256, 4)
(tccg/gett.py:1008-1059, tccg/gett.py:689-698).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time
from pathlib import Path
from typing import Dict, List, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = _HERE / "_build" / "libavx2gett.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_HERE), "_build/libavx2gett.so"], check=True)
    return _LIB


def _load():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            build()
        lib = ctypes.CDLL(str(_LIB))
        P = ctypes.c_void_p
        for name, T in (("avx2_gett_f32", ctypes.c_float), ("avx2_gett_f64", ctypes.c_double)):
            fn = getattr(lib, name)
            fn.argtypes = [P, P, P, T, T, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64] + [P] * 6 + \
                [ctypes.c_int] * 4
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


def possible_values(block: int, must: int, target: int, maxn: int) -> List[int]:
    out = []
    a = target // block
    while len(out) < maxn and a >= 1:
        while must % (a * block) != 0 and a > 1:
            a -= 1
        if a * block > 0 and must % (a * block) == 0:
            out.append(a * block)
        a -= 1
    return out


def _strides(labels, sizes):
    st, s = {}, 1
    for l in labels:
        st[l] = s
        s *= sizes[l]
    return st


def _table(order: Sequence[str], sizes: Dict[str, int], stride: Dict[str, int]) -> np.ndarray:
    """int64 offsets over the column-major linearisation of `order`."""
    n = int(np.prod([sizes[l] for l in order])) if order else 1
    p = np.arange(n, dtype=np.int64)
    off = np.zeros(n, dtype=np.int64)
    for l in order:
        e = sizes[l]
        off += (p % e) * stride.get(l, 0)
        p //= e
    return off


class Avx2Gett:
    """The restated AVX2 GETT for one contraction (c, a, b label lists)."""

    def __init__(self, c, a, b, sizes: Dict[str, int], dtype=np.float32):
        self.c, self.a, self.b, self.sizes, self.dtype = list(c), list(a), list(b), dict(sizes), dtype
        holder_is_a = c[0] in a
        X, Y = (a, b) if holder_is_a else (b, a)
        m = [l for l in c if l in X and l not in Y]
        n = [l for l in c if l in Y and l not in X]
        k = [l for l in a if l in b and l not in c]
        self.swapped = not holder_is_a
        sx, sy, sc = _strides(X, sizes), _strides(Y, sizes), _strides(c, sizes)
        self.M = int(np.prod([sizes[l] for l in m])) if m else 1
        self.N = int(np.prod([sizes[l] for l in n])) if n else 1
        self.K = int(np.prod([sizes[l] for l in k])) if k else 1
        self.tabs = [_table(m, sizes, sx), _table(k, sizes, sx), _table(n, sizes, sy), _table(k, sizes, sy),
                     _table(m, sizes, sc), _table(n, sizes, sc)]
        dbl = dtype == np.float64
        mr, nr = (12, 4) if dbl else (24, 4)
        self.mr, self.nr = mr, nr
        mcs = possible_values(mr, self.M, 96 if dbl else 192, 3)
        self.mc = mcs[0] if mcs else 0
        self.nc = self.N if self.N < 4600 else (possible_values(nr, self.N, 4600, 3) or [0])[0]
        kcs = possible_values(1, self.K, 256, 4)
        self.kc = kcs[0] if kcs else 0

    def supported(self) -> bool:
        return (self.mc > 0 and self.nc > 0 and self.kc > 0 and self.M % self.mr == 0 and self.N % self.nr == 0)

    def flops(self) -> float:
        return 2.0 * self.M * self.N * self.K

    def __call__(self, A, B, C, alpha=1.0, beta=0.0, threads: int = 0):
        if not self.supported():
            raise ValueError("extents outside the reference GETT's limits (M %% %d, N %% %d)" % (self.mr, self.nr))
        lib = _load()
        A = np.ascontiguousarray(A, dtype=self.dtype)
        B = np.ascontiguousarray(B, dtype=self.dtype)
        assert C.dtype == self.dtype and C.flags.c_contiguous
        X, Y = (B, A) if self.swapped else (A, B)
        ptr = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        fn = lib.avx2_gett_f64 if self.dtype == np.float64 else lib.avx2_gett_f32
        rc = fn(ptr(X), ptr(Y), ptr(C), alpha, beta, self.M, self.N, self.K, *[ptr(t) for t in self.tabs],
                self.mc, self.nc, self.kc, threads)
        if rc != 0:
            raise RuntimeError("avx2_gett rejected the blocking (rc=%d)" % rc)
        return C


def physical_cores() -> int:
    try:
        import psutil
        n = psutil.cpu_count(logical=False)
        if n:
            return int(n)
    except Exception:
        pass
    return os.cpu_count() or 1


def bench_workload(name: str, seconds: float = 15.0, log=print) -> dict:
    """Time the restated AVX2 GETT on the named workload: whole contractions,
    minimum over repetitions, on all physical cores and on one core."""
    import re
    from tccg_amd.frontend import parse_tccg  # parser only: labels and sizes
    from tccg_amd.workloads import get
    C_, A_, B_, alpha, beta = parse_tccg(get(name))
    c, a, b = ([i.label for i in T.indices] for T in (C_, A_, B_))
    sizes = {i.label: i.size for T in (C_, A_, B_) for i in T.indices}
    g = Avx2Gett(c, a, b, sizes)
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(np.float32)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(np.float32)
    C = np.zeros([sizes[l] for l in reversed(c)], dtype=np.float32)
    cores = physical_cores()
    out = {}
    budget = {cores: 0.6 * seconds, 1: 0.4 * seconds}
    for nt in (cores, 1):
        best, t_end, reps = float("inf"), time.perf_counter() + budget[nt], 0
        while reps < 1 or (time.perf_counter() < t_end and reps < 50):
            t0 = time.perf_counter()
            g(A, B, C, alpha, beta, threads=nt)
            best = min(best, time.perf_counter() - t0)
            reps += 1
        out[nt] = (g.flops() / best / 1e9, reps, best)
        log("cpu baseline (restated AVX2 GETT) %s: %d threads %.1f GFLOP/s (min of %d runs, %.3f s)"
            % (name, nt, out[nt][0], reps, best))
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"model name\s*:\s*(.*)", f.read())
            model = m.group(1).strip() if m else ""
    except OSError:
        pass
    return {"value": round(out[cores][0], 1), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "sample": "whole %s contraction (mc=%d nc=%d kc=%d), min over %d runs; 1 core: %.1f GFLOP/s (%d runs); "
                      "CPU %s" % (name, g.mc, g.nc, g.kc, out[cores][1], out[1][0], out[1][1], model),
            "one_core_value": round(out[1][0], 1)}
