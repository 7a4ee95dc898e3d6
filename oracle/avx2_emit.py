"""Per-contraction AVX2 GETT source emitter -- TEST INFRASTRUCTURE ONLY (the
CPU baseline and an independent CPU check of the GPU results).

The reference generates one C++ function per GETT candidate with the
extents, blockings and loop order baked in, its micro-kernel written out
instruction by instruction by the avx2 emitter (tccg/arch.py:57-151:
setzero, vector load, broadcast, fmadd, store per register), the k loop
unrolled by two (tccg/gett.py:145-227), the macro-kernel and the loop
variant's driver around it (tccg/gett.py:229-307, 1237-1400), entry
`void gett(T* A, T* B, T* C, const T alpha, const T beta)`
(tccg/gett.py:342-344).  This module emits the same structure: the
micro-kernel as straight-line intrinsics over named accumulator registers
(MR/W x NR of them: 3 x 4 for float, 3 x 4 for double), the blocking of the
chosen candidate (oracle.avx2_model) as compile-time constants, and the
index tables of the contraction baked in as constant arrays (the reference
bakes strides into HPTT transposition plans; a constant gather is its
counterpart here).  Arithmetic order matches oracle/avx2_gett.cpp, the
table-driven restatement, so both give the same bits for the same
candidate.
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import subprocess
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np

from .avx2_gett import Avx2Gett

_HERE = Path(__file__).resolve().parent
EMIT_DIR = _HERE / "_build" / "emitted"
CXXFLAGS = ["-O3", "-fPIC", "-std=c++17", "-mavx2", "-mfma", "-fopenmp", "-shared"]


class _Isa:
    """The avx2 instruction forms (tccg/arch.py:57-151) for one type."""

    def __init__(self, dbl: bool):
        self.T = "double" if dbl else "float"
        self.V = "__m256d" if dbl else "__m256"
        s = "pd" if dbl else "ps"
        self.W = 4 if dbl else 8
        self.MR, self.NR = (12, 4) if dbl else (24, 4)
        self.zero = "_mm256_setzero_%s()" % s
        self.load = "_mm256_loadu_%s(%%s)" % s
        self.bcast = "_mm256_broadcast_%s(%%s)" % ("sd" if dbl else "ss")
        self.fma = "_mm256_fmadd_%s(%%s, %%s, %%s)" % s
        self.store = "_mm256_storeu_%s(%%s, %%s);" % s
        self.add = "_mm256_add_%s(%%s, %%s)" % s
        self.set1 = "_mm256_set1_%s(%%s)" % s


def _array(name: str, values: np.ndarray) -> str:
    body = ",".join(str(int(v)) for v in values)
    return "static const int64_t %s[%d] = {%s};\n" % (name, max(1, values.size), body or "0")


def emit_micro_kernel(isa: _Isa, kc: int) -> str:
    """MR x NR micro-kernel over one KC block, straight-line per k and
    unrolled by two; the tile goes to `out` column by column (MR rows)."""
    ra = isa.MR // isa.W
    L = ["static inline void micro_kernel(const %s* __restrict__ Ap, const %s* __restrict__ Bp, "
         "%s* __restrict__ out) {" % (isa.T, isa.T, isa.T)]
    for j in range(isa.NR):
        for i in range(ra):
            L.append("  %s c_%d_%d = %s;" % (isa.V, j, i, isa.zero))

    def body(step: int):
        out = []
        for i in range(ra):
            out.append("    const %s a_%d_%d = %s;" % (isa.V, step, i, isa.load % ("Ap + %d" % (step * isa.MR + i * isa.W))))
        for j in range(isa.NR):
            out.append("    const %s b_%d_%d = %s;" % (isa.V, step, j, isa.bcast % ("Bp + %d" % (step * isa.NR + j))))
            for i in range(ra):
                out.append("    c_%d_%d = %s;" % (j, i, isa.fma % ("a_%d_%d" % (step, i), "b_%d_%d" % (step, j),
                                                                 "c_%d_%d" % (j, i))))
        return out

    L.append("  for (int k = 0; k < %d; ++k) {" % (kc // 2))
    L += body(0) + body(1)
    L.append("    Ap += %d;" % (2 * isa.MR))
    L.append("    Bp += %d;" % (2 * isa.NR))
    L.append("  }")
    if kc % 2:
        L.append("  {")
        L += body(0)
        L.append("  }")
    for j in range(isa.NR):
        for i in range(ra):
            L.append("  " + isa.store % ("out + %d" % (j * isa.MR + i * isa.W), "c_%d_%d" % (j, i)))
    L.append("}")
    return "\n".join(L) + "\n"


_DRIVER = r"""
// Pack A^ (rows m0 .. m0+rows-1 at k0) into MR-slivers A~[rows/MR][KC][MR], times scale.
static inline void pack_A(const T* A, int64_t m0, int64_t k0, int rows, T scale, T* dst0) {
  for (int64_t s = 0; s < rows / MR; ++s) {
    T* dst = dst0 + (size_t)s * KC * MR;
    const int64_t* am = kAm + m0 + s * MR;
    for (int k = 0; k < KC; ++k) {
      const int64_t ak = kAk[k0 + k];
      for (int i = 0; i < MR; ++i) dst[k * MR + i] = scale * A[am[i] + ak];
    }
  }
}

// Pack sliver s of B^ (columns n0 + s*NR .. at k0) into B~[NC/NR][KC][NR], times scale.
static inline void pack_B_sliver(const T* B, int64_t n0, int64_t k0, int64_t s, T scale, T* dst0) {
  T* dst = dst0 + (size_t)s * KC * NR;
  for (int k = 0; k < KC; ++k) {
    const int64_t bk = kBk[k0 + k];
    for (int j = 0; j < NR; ++j) dst[k * NR + j] = scale * B[kBn[n0 + s * NR + j] + bk];
  }
}

// One micro-tile: C = tile (+ beta C on the first K block, + C on later ones).
static inline void micro_update(const T* Ap, const T* Bp, int64_t mb, int64_t nb, bool first, T beta, T* C) {
  alignas(32) T tile[MR * NR];
  micro_kernel(Ap, Bp, tile);
  for (int j = 0; j < NR; ++j) {
    const int64_t cn = kCn[nb + j];
    T* col = tile + j * MR;
    if (kCmContig[mb / MR]) {
      T* c = C + kCm[mb] + cn;
      for (int i = 0; i < MR; i += W) {
        V v = LOAD(col + i);
        if (!first) v = ADD(v, LOAD(c + i));
        else if (beta != T(0)) v = FMA(SET1(beta), LOAD(c + i), v);
        STORE(c + i, v);
      }
    } else {
      for (int i = 0; i < MR; ++i) {
        T* c = C + kCm[mb + i] + cn;
        if (!first) *c += col[i];
        else if (beta != T(0)) *c = col[i] + beta * *c;
        else *c = col[i];
      }
    }
  }
}

int g_threads = 0;

static T* alloc(size_t n) { return (T*)aligned_alloc(64, (n * sizeof(T) + 63) & ~(size_t)63); }

static void run(const T* X, const T* Y, T* C, T alpha, T beta) {
  if (g_threads > 0) omp_set_num_threads(g_threads);
  const int nth = g_threads > 0 ? g_threads : omp_get_max_threads();
#if VARIANT == 1
  // m outer: A~ packed once per (m0, k0) scaled by alpha, B~ per thread
  T* At = alloc((size_t)MC * KC);
  T* Bt = alloc((size_t)nth * NC * KC);
  for (int64_t m0 = 0; m0 < M; m0 += MC) {
    for (int64_t k0 = 0; k0 < K; k0 += KC) {
      const bool first = (k0 == 0);
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < MC / MR; ++s) pack_A(X, m0 + s * MR, k0, MR, alpha, At + s * KC * MR);
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t n0 = 0; n0 < N; n0 += NC) {
        T* Bt_ = Bt + (size_t)omp_get_thread_num() * NC * KC;
        for (int64_t s = 0; s < NC / NR; ++s) pack_B_sliver(Y, n0, k0, s, T(1), Bt_);
        for (int64_t im = 0; im < MC / MR; ++im)
          for (int64_t jn = 0; jn < NC / NR; ++jn)
            micro_update(At + (size_t)im * KC * MR, Bt_ + (size_t)jn * KC * NR, m0 + im * MR, n0 + jn * NR, first,
                         beta, C);
      }
    }
  }
#else
  // n outer: B~ packed once per (n0, k0) scaled by alpha, A~ per thread
  T* Bt = alloc((size_t)NC * KC);
  T* At = alloc((size_t)nth * MC * KC);
  for (int64_t n0 = 0; n0 < N; n0 += NC) {
    for (int64_t k0 = 0; k0 < K; k0 += KC) {
      const bool first = (k0 == 0);
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < NC / NR; ++s) pack_B_sliver(Y, n0, k0, s, alpha, Bt);
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t m0 = 0; m0 < M; m0 += MC) {
        T* At_ = At + (size_t)omp_get_thread_num() * MC * KC;
        pack_A(X, m0, k0, MC, T(1), At_);
        for (int64_t jn = 0; jn < NC / NR; ++jn)
          for (int64_t im = 0; im < MC / MR; ++im)
            micro_update(At_ + (size_t)im * KC * MR, Bt + (size_t)jn * KC * NR, m0 + im * MR, n0 + jn * NR, first,
                         beta, C);
      }
    }
  }
#endif
  free(At);
  free(Bt);
}

}  // namespace

extern "C" {

// The reference's generated entry point (tccg/gett.py:342-344): A and B in
// the contraction's order.
void gett(T* A, T* B, T* C, const T alpha, const T beta) {
  if (SWAPPED) run(B, A, C, alpha, beta);
  else run(A, B, C, alpha, beta);
}

void gett_set_threads(int n) { g_threads = n; }

}  // extern "C"
"""


def emit(g: Avx2Gett, label: str = "") -> str:
    """C++ source of the GETT for `g`'s contraction with `g`'s candidate."""
    if not g.supported():
        raise ValueError("the candidate's blocking does not divide the extents")
    dbl = g.dtype == np.float64
    isa = _Isa(dbl)
    Am, Ak, Bn, Bk, Cm, Cn = g.tabs
    contig = [bool(np.all(Cm[s * isa.MR:(s + 1) * isa.MR] == Cm[s * isa.MR] + np.arange(isa.MR)))
              for s in range(g.M // isa.MR)]
    cand = g.candidate.name() if g.candidate is not None else "variant %d" % g.variant
    head = [
        "// Generated by oracle/avx2_emit.py -- test infrastructure (CPU baseline).",
        "// Contraction: %s" % label,
        "// Candidate: %s" % cand,
        "#include <immintrin.h>",
        "#include <omp.h>",
        "#include <stdint.h>",
        "#include <stdlib.h>",
        "",
        "namespace {",
        "",
        "using T = %s;" % isa.T,
        "using V = %s;" % isa.V,
        "#define LOAD(p) %s" % (isa.load % "p"),
        "#define STORE(p, v) %s" % (isa.store % ("p", "v")).rstrip(";"),
        "#define ADD(a, b) %s" % (isa.add % ("a", "b")),
        "#define FMA(a, b, c) %s" % (isa.fma % ("a", "b", "c")),
        "#define SET1(x) %s" % (isa.set1 % "x"),
        "#define VARIANT %d" % g.variant,
        "#define SWAPPED %d" % (1 if g.swapped else 0),
        "constexpr int64_t M = %d, N = %d, K = %d;" % (g.M, g.N, g.K),
        "constexpr int MC = %d, NC = %d, KC = %d, MR = %d, NR = %d, W = %d;" % (g.mc, g.nc, g.kc, isa.MR, isa.NR,
                                                                             isa.W),
        "",
    ]
    tables = "".join(_array(n, t) for n, t in (("kAm", Am), ("kAk", Ak), ("kBn", Bn), ("kBk", Bk), ("kCm", Cm),
                                              ("kCn", Cn)))
    tables += "static const bool kCmContig[%d] = {%s};\n" % (len(contig), ",".join("1" if c else "0" for c in contig))
    return "\n".join(head) + tables + "\n" + emit_micro_kernel(isa, g.kc) + _DRIVER


def _label(c, a, b) -> str:
    return "C[%s] = A[%s] * B[%s]" % (",".join(c), ",".join(a), ",".join(b))


def compile_source(src: str, out_dir: Optional[Path] = None) -> Path:
    """Compile (once per distinct source) into a shared library with g++."""
    out_dir = Path(out_dir) if out_dir is not None else EMIT_DIR
    out_dir.mkdir(parents=True, exist_ok=True)
    digest = hashlib.sha256((src + " ".join(CXXFLAGS)).encode()).hexdigest()[:20]
    so = out_dir / ("gett_%s.so" % digest)
    if not so.exists():
        cpp = out_dir / ("gett_%s.cpp" % digest)
        cpp.write_text(src)
        tmp = so.with_suffix(".so.tmp%d" % os.getpid())
        try:
            subprocess.run(["g++"] + CXXFLAGS + ["-o", str(tmp), str(cpp)], check=True)
            tmp.rename(so)
        finally:
            cpp.unlink()   # the source is regenerated from the contraction on demand
            if tmp.exists():
                tmp.unlink()
    return so


class EmittedGett:
    """The emitted and compiled GETT of one contraction and candidate, called
    through its `gett(A, B, C, alpha, beta)` entry (A and B in the
    contraction's order, TCCG layout)."""

    def __init__(self, c, a, b, sizes: Dict[str, int], candidate, dtype=np.float32, out_dir=None):
        self.g = Avx2Gett(c, a, b, sizes, dtype, candidate)
        self.dtype = dtype
        self.path = compile_source(emit(self.g, _label(c, a, b)), out_dir)
        self.lib = ctypes.CDLL(str(self.path))
        T = ctypes.c_double if dtype == np.float64 else ctypes.c_float
        P = ctypes.c_void_p
        self.lib.gett.argtypes = [P, P, P, T, T]
        self.lib.gett.restype = None
        self.lib.gett_set_threads.argtypes = [ctypes.c_int]

    def flops(self) -> float:
        return self.g.flops()

    def __call__(self, A, B, C, alpha=1.0, beta=0.0, threads: int = 0):
        A = np.ascontiguousarray(A, dtype=self.dtype)
        B = np.ascontiguousarray(B, dtype=self.dtype)
        assert C.dtype == self.dtype and C.flags.c_contiguous
        ptr = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        self.lib.gett_set_threads(int(threads))
        self.lib.gett(ptr(A), ptr(B), ptr(C), alpha, beta)
        return C


def emit_for(c: Sequence[str], a: Sequence[str], b: Sequence[str], sizes: Dict[str, int], top: int = 3,
             dtype=np.float32):
    """EmittedGett objects for the restated planner's `top` candidates."""
    from .avx2_model import candidates_for
    ranked = candidates_for(list(c), list(a), list(b), sizes, "double" if dtype == np.float64 else "float", 0.0, 16)
    out = []
    for cand in ranked:
        g = Avx2Gett(c, a, b, sizes, dtype, cand)
        if g.supported():
            out.append(EmittedGett(c, a, b, sizes, cand, dtype))
        if len(out) >= top:
            break
    return out


def workload_labels(name: str):
    """(c, a, b, sizes) of a named workload (parser only)."""
    from tccg_amd.frontend import parse_tccg
    from tccg_amd.workloads import get
    C_, A_, B_, _, _ = parse_tccg(get(name))
    c, a, b = ([i.label for i in T.indices] for T in (C_, A_, B_))
    sizes = {i.label: i.size for T in (C_, A_, B_) for i in T.indices}
    return c, a, b, sizes


def prebuild(workloads: Sequence[str], top: int = 3, jobs: int = 8) -> int:
    """Emit and compile the planner's `top` candidates of each workload (FP32)
    in parallel, so that the CPU baseline finds them built; returns the
    number of libraries."""
    from concurrent.futures import ThreadPoolExecutor
    from .avx2_model import candidates_for
    srcs = []
    for wl in workloads:
        c, a, b, sizes = workload_labels(wl)
        n = 0
        for cand in candidates_for(c, a, b, sizes, "float", 0.0, 16):
            g = Avx2Gett(c, a, b, sizes, np.float32, cand)
            if g.supported():
                srcs.append(emit(g, _label(c, a, b)))
                n += 1
            if n >= top:
                break
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        return len(list(ex.map(compile_source, srcs)))
