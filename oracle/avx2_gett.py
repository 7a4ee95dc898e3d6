"""ctypes driver of avx2_gett.cpp, the restated AVX2 GETT -- test infrastructure
and CPU baseline only.

Classification (m/n/k sets, A/B swap so that the internal A holds C's
stride-one index) follows tccg/tccg_util.py:212-296, tccg/gett.py:64-89.
The loop variant and block sizes (mr=24/nr=4 float, 12/4 double; mc, nc, kc)
come from a candidate of the restated AVX2 planner (oracle/avx2_model.py);
by default its best-ranked one, the reference's policy
(tccg/gett.py:961-1150).  The restated kernel blocks the m and n sets as
ranges of their linearisation in C's order (the reference blocks the split
sub-index sets that the candidate's name records); a block holds mc (nc)
positions either way.
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
                [ctypes.c_int] * 5
            fn.restype = ctypes.c_int
        _lib = lib
    return _lib


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
    """The restated AVX2 GETT for one contraction (c, a, b label lists) with
    the blocking of `candidate` (an oracle.avx2_model.Avx2Candidate; default:
    the planner's best-ranked candidate)."""

    def __init__(self, c, a, b, sizes: Dict[str, int], dtype=np.float32, candidate=None):
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
        self.mr, self.nr = (12, 4) if dbl else (24, 4)
        self.variant, self.mc, self.nc, self.kc = 0, 0, 0, 0
        self.candidate = candidate
        if candidate is None and self.M % self.mr == 0 and self.N % self.nr == 0:
            from .avx2_model import candidates_for
            ranked = candidates_for(c, a, b, sizes, "double" if dbl else "float", 0.0, 1)
            candidate = ranked[0] if ranked else None
            self.candidate = candidate
        if candidate is not None:
            self.variant, self.mc, self.nc, self.kc = candidate.variant, candidate.mc, candidate.nc, candidate.kc

    def supported(self) -> bool:
        return (self.mc > 0 and self.nc > 0 and self.kc > 0 and self.M % self.mr == 0 and self.N % self.nr == 0
                and self.M % self.mc == 0 and self.N % self.nc == 0 and self.K % self.kc == 0
                and self.mc % self.mr == 0 and self.nc % self.nr == 0)

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
                self.variant, self.mc, self.nc, self.kc, threads)
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


def cpu_model() -> str:
    import re
    try:
        with open("/proc/cpuinfo") as f:
            m = re.search(r"model name\s*:\s*(.*)", f.read())
            return m.group(1).strip() if m else ""
    except OSError:
        return ""


def _time(fn, budget: float, max_reps: int) -> tuple:
    best, t_end, reps = float("inf"), time.perf_counter() + budget, 0
    while reps < 1 or (time.perf_counter() < t_end and reps < max_reps):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
        reps += 1
    return best, reps


def bench_workload(name: str, seconds: float = 5.0, top: int = 3, log=print) -> dict:
    """The reference's measure-and-select on the host (tccg/tccg.py:987-1058,
    timing by min over repetitions, tccg/tccg.py:725-737): the restated
    planner's `top` best-ranked candidates, each emitted as its own
    per-contraction source and compiled (oracle/avx2_emit.py, as the
    reference generates and compiles one function per candidate), timed on
    all physical cores, the fastest kept and also timed on one core; its
    result checked against the FP64 einsum oracle (relative Frobenius error
    <= 1e-5) and against the table-driven restatement (same bits)."""
    from .avx2_emit import EmittedGett, workload_labels
    from .avx2_model import candidates_for
    from .einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.frontend import parse_tccg  # parser only: alpha and beta
    from tccg_amd.workloads import get
    _, _, _, alpha, beta = parse_tccg(get(name))
    c, a, b, sizes = workload_labels(name)
    ranked = candidates_for(c, a, b, sizes, "float", beta, 16)
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(np.float32)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(np.float32)
    C = np.zeros([sizes[l] for l in reversed(c)], dtype=np.float32)
    cores = physical_cores()
    timed = []
    for cand in ranked:
        if not Avx2Gett(c, a, b, sizes, candidate=cand).supported():
            continue
        g = EmittedGett(c, a, b, sizes, cand)
        t, reps = _time(lambda: g(A, B, C, alpha, beta, threads=cores), seconds * 0.5 / top, 20)
        timed.append((t, reps, g, cand))
        log("cpu baseline %s %s: estimated %.1f GFLOP/s (one Haswell core), %d threads %.1f GFLOP/s"
            % (name, cand.name()[:40], cand.estimate, cores, g.flops() / t / 1e9))
        if len(timed) >= top:
            break
    if not timed:
        raise ValueError("no candidate of the restated planner runs %s" % name)
    t_all, reps_all, g, cand = min(timed, key=lambda x: x[0])
    t_one, reps_one = _time(lambda: g(A, B, C, alpha, beta, threads=1), seconds * 0.5, 10)
    C[...] = 0
    g(A, B, C, alpha, beta, threads=cores)
    err = rel_frobenius(C, contract_ref(c, a, b, A, B, None, alpha, 0.0))
    if not err <= 1e-5:
        raise RuntimeError("emitted AVX2 GETT misses the einsum oracle on %s: %.3e" % (name, err))
    C2 = np.zeros_like(C)
    Avx2Gett(c, a, b, sizes, candidate=cand)(A, B, C2, alpha, beta, threads=cores)
    if not np.array_equal(C, C2):
        raise RuntimeError("emitted and table-driven AVX2 GETT differ on %s" % name)
    gf_all, gf_one = g.flops() / t_all / 1e9, g.flops() / t_one / 1e9
    log("cpu baseline %s: %s, %d threads %.1f GFLOP/s, 1 thread %.1f GFLOP/s, rel. err %.1e"
        % (name, cand.name()[:60], cores, gf_all, gf_one, err))
    return {"workload": name, "candidate": cand.name(), "code": "emitted per contraction",
            "estimated_gflops_one_haswell_core": round(cand.estimate, 2),
            "gflops": round(gf_all, 1), "gflops_one_core": round(gf_one, 1), "cores": cores,
            "candidates_timed": len(timed), "reps": reps_all, "reps_one_core": reps_one, "rel_err_vs_fp64": err}
