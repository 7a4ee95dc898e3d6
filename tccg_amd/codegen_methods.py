"""Per-contraction libraries of the TTGT and LoG methods.

When the CLI finds TTGT or LoG faster than the best GETT, the reference
emits that method's code instead of the GETT (emitFinalCode,
tccg/tccg.py:158-186): `ttgemmt.cpp` with
`int ttgemmt_<name>(const T* A, const T* B, T* C, T alpha, T beta, T* work_)`
(tccg/ttgemmt.py:210-214), or `loopOverGemm.cpp` with
`int <LoG name>(const T* A, const T* B, T* C, T alpha, T beta)`
(tccg/gemm_loop.py:88-96).  Those call HPTT and BLAS; the sources emitted
here call this backend's library (libtccg_gfx950.so: the transposition
kernels, the packed or gather GETT for TTGT's GEMM, the grouped launch for
LoG), with the plans' tables baked in as device data, and export the
reference's method entry point beside the uniform boundary
(`gett`, `gett_stream`, `gett_info`, `gett_contraction`, `gett_last_error`),
so a caller of either keeps working.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import build as _build
from .plan import F32, F64, PREC_NAMES

_CTYPE = {F32: "float", F64: "double"}
_ALIGN = 256


def _words(name: str, ctype: str, values: Sequence[int]) -> str:
    vals = list(values) or [0]
    lines = []
    for i in range(0, len(vals), 24):
        lines.append(",".join(str(int(v)) for v in vals[i:i + 24]))
    return "__device__ __attribute__((used)) %s %s[%d] = {\n%s};\n" % (ctype, name, len(vals), ",\n".join(lines))


def _host_array(name: str, ctype: str, values: Sequence[int]) -> str:
    vals = list(values) or [0]
    lines = []
    for i in range(0, len(vals), 24):
        lines.append(",".join(str(int(v)) for v in vals[i:i + 24]))
    return "static const %s %s[%d] = {\n%s};\n" % (ctype, name, len(vals), ",\n".join(lines))


class _Blob:
    """uint32 tables concatenated, each 16-byte aligned; offsets in bytes."""

    def __init__(self):
        self.words: List[int] = []
        self.ids: Dict[int, int] = {}
        self.keep: List[np.ndarray] = []   # holds every added array, so no id() is reused

    def add(self, arr: np.ndarray) -> int:
        key = id(arr)
        if key in self.ids:
            return self.ids[key]
        self.keep.append(arr)
        off = 4 * len(self.words)
        self.words.extend(np.ascontiguousarray(arr).astype(np.uint32).view(np.int32).tolist())
        self.words.extend([0] * ((-len(self.words)) % 4))
        self.ids[key] = off
        return off


_PRELUDE = r"""#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "tccg_gfx950.h"
"""

_STATE = r"""
constexpr int kMaxDevices = 64;
thread_local std::string g_err;
std::mutex g_init;   // serialises first-call set-up (allocations included)

int fail(int code, const std::string& m) { g_err = m; return code ? code : 1; }
int lib_fail(int code, const char* what) { return fail(code, std::string(what) + ": " + tccg_last_error()); }

// Device address of a baked-in table, after checking that the code object
// holds it with the expected size.
int symbol(const void*& dst, const void* sym, size_t bytes) {
  void* p = nullptr;
  size_t n = 0;
  hipError_t e = hipGetSymbolSize(&n, sym);
  if (e != hipSuccess) return fail((int)e, std::string("hipGetSymbolSize: ") + hipGetErrorString(e));
  if (n != bytes) return fail(1, "device table has an unexpected size");
  e = hipGetSymbolAddress(&p, sym);
  if (e != hipSuccess) return fail((int)e, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  if (!p) return fail(1, "device table address is NULL");
  dst = p;
  return 0;
}

int device_index(int* dev) {
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) return fail((int)e, "hipGetDevice failed");
  if (*dev < 0 || *dev >= kMaxDevices) return fail(1, "device index out of range");
  return 0;
}
"""


def _tplan(blob: _Blob, tp) -> Dict[str, int]:
    return dict(inX=blob.add(tp.inX), outX=blob.add(tp.outX), inY=blob.add(tp.inY), outY=blob.add(tp.outY),
                inZ=blob.add(tp.inZ), outZ=blob.add(tp.outZ), Xp=tp.Xp, Yp=tp.Yp, Z=tp.nZ, tile=tp.tile, vec=tp.vec)


def _tplan_code(var: str, d: Dict[str, int]) -> str:
    return ("  tccg_transpose_plan %(v)s;\n"
            "  %(v)s.inX = W(%(inX)d); %(v)s.outX = W(%(outX)d); %(v)s.inY = W(%(inY)d); %(v)s.outY = W(%(outY)d);\n"
            "  %(v)s.inZ = W(%(inZ)d); %(v)s.outZ = W(%(outZ)d);\n"
            "  %(v)s.Xp = %(Xp)d; %(v)s.Yp = %(Yp)d; %(v)s.Z = %(Z)d; %(v)s.tile = %(tile)d; %(v)s.vec = %(vec)d;\n"
            % dict(v=var, **d))


def _info(contraction, prec: int, tile: int, BM: int, BN: int, BK: int, modeA: int, modeB: int) -> List[int]:
    pb = contraction.problem
    sizeA, sizeB = (pb.opB.size(), pb.opA.size()) if pb.swapped else (pb.opA.size(), pb.opB.size())
    return [pb.M, pb.N, pb.K, prec, tile, BM, BN, BK, modeA, modeB, sizeA, sizeB, pb.C.size()]


def _tail(T: str, info: List[int], desc: str, method_entry: str, method_call: str) -> str:
    return """
void gett(%(T)s* A, %(T)s* B, %(T)s* C, const %(T)s alpha, const %(T)s beta) {
  (void)gett_stream(A, B, C, alpha, beta, nullptr);
}

%(entry)s {
  return %(call)s;
}

int gett_info(int64_t* out, int n) {
  const int64_t v[%(ni)d] = {%(info)s};
  for (int i = 0; i < n && i < %(ni)d; ++i) out[i] = v[i];
  return %(ni)d;
}

const char* gett_contraction(void) { return "%(desc)s"; }

const char* gett_last_error(void) { return g_err.c_str(); }

}  // extern "C"
""" % dict(T=T, entry=method_entry, call=method_call, ni=len(info), info=",".join(str(int(x)) for x in info),
           desc=desc)


def _desc(text: str) -> str:
    return text.strip().replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n")


def emit_ttgt_source(contraction, cand, prec: int, gemm_plan, contraction_text: str) -> str:
    """HIP source of the TTGT method for `cand` (tccg_amd.ttgt): the
    transpositions of the candidate, its GEMM by `gemm_plan` (this backend's
    GETT of the transposed operands), temporaries in one workspace."""
    from .contraction import Contraction
    from .transpose import plan_transpose
    if prec not in _CTYPE:
        raise ValueError("TTGT libraries are emitted for FP32 and FP64")
    T = _CTYPE[prec]
    esz = 8 if prec == F64 else 4
    pb = contraction.problem
    sizes = {i.label: i.size for X in (pb.opA, pb.opB, pb.C) for i in X.indices}
    X, Y = (pb.opB, pb.opA) if cand.swap else (pb.opA, pb.opB)
    blob = _Blob()
    offs, total = {}, 0
    for key, need, labels in (("A", cand.transA, cand.idxA), ("B", cand.transB, cand.idxB),
                              ("C", cand.transC, cand.idxC)):
        if need:
            offs[key] = total
            total += -(-int(np.prod([sizes[l] for l in labels])) * esz // _ALIGN) * _ALIGN
    tA = _tplan(blob, plan_transpose(X.labels(), cand.idxA, sizes, elem_bytes=esz)) if cand.transA else None
    tB = _tplan(blob, plan_transpose(Y.labels(), cand.idxB, sizes, elem_bytes=esz)) if cand.transB else None
    tC = _tplan(blob, plan_transpose(cand.idxC, pb.C.labels(), sizes, elem_bytes=esz)) if cand.transC else None
    gemm = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes)
    g = gemm_plan
    if g.prec != prec:
        raise ValueError("the TTGT GEMM plan must be a %s plan" % PREC_NAMES[prec])
    gt = {k: blob.add(getattr(g, k)) for k in ("offAm", "offAk", "offBn", "offBk", "offCm", "offCn", "ordA", "ordB")}
    ws_gemm = g.workspace_bytes() if g.packed else 0
    crun = (g.crun[0], 0) if g.walk == 4 else (g.crun if not g.grid or g.walk == 5 else (0, 0))
    fn = {F32: "f32", F64: "f64"}[prec]
    gemm_call = ("tccg_gett_packed_%s(gA, gB, out, alpha_, beta_, &gp, st->ws, %d, s)" % (fn, ws_gemm) if g.packed
                 else "tccg_gett_%s(gA, gB, out, alpha_, beta_, &gp, s)" % fn)
    name = "ttgemmt_" + cand.name
    src = ["// Generated by tccg_amd (gfx950 target): the TTGT method, candidate %s; do not edit.\n" % name]
    for line in contraction_text.strip().splitlines():
        src.append("//   %s\n" % line)
    src.append("// GEMM plan: %s\n" % g.describe())
    src.append(_PRELUDE)
    src.append(_words("ttgt_tables", "int32_t", blob.words))
    src.append("\nnamespace {\n")
    src.append(_STATE)
    src.append("""
struct State {
  const unsigned char* tables;
  void* tmp;   // transposed temporaries (%(tmp)d bytes)
  void* ws;    // the GEMM's packed-operand workspace (%(ws)d bytes)
  std::atomic<bool> ready;
};
State g_state[kMaxDevices];

int state_for_current_device(State** out) {
  int dev = 0;
  if (int rc = device_index(&dev)) return rc;
  State& t = g_state[dev];
  if (!t.ready.load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lock(g_init);
    if (!t.ready.load(std::memory_order_relaxed)) {
      const void* p = nullptr;
      if (int rc = symbol(p, HIP_SYMBOL(ttgt_tables), %(nb)d)) return rc;
      t.tables = static_cast<const unsigned char*>(p);
      hipError_t e = hipSuccess;
      if (%(tmp)d > 0 && (e = hipMalloc(&t.tmp, %(tmp)d)) != hipSuccess)
        return fail((int)e, "hipMalloc of the TTGT temporaries failed");
      if (%(ws)d > 0 && (e = hipMalloc(&t.ws, %(ws)d)) != hipSuccess)
        return fail((int)e, "hipMalloc of the GEMM workspace failed");
      t.ready.store(true, std::memory_order_release);
    }
  }
  *out = &t;
  return 0;
}

}  // namespace

extern "C" {

int gett_stream(const %(T)s* A, const %(T)s* B, %(T)s* C, %(T)s alpha, %(T)s beta, void* stream) {
  if (!A || !B || !C) return fail(1, "tensor pointer is NULL");
  State* st = nullptr;
  if (int rc = state_for_current_device(&st)) return rc;
  hipStream_t s = (hipStream_t)stream;
  auto W = [&](int off) { return reinterpret_cast<const uint32_t*>(st->tables + off); };
  unsigned char* tmp = static_cast<unsigned char*>(st->tmp);
  (void)tmp;
  const %(T)s* opA = %(swapped)d ? B : A;
  const %(T)s* opB = %(swapped)d ? A : B;
  const %(T)s* X = %(xb)d ? opB : opA;
  const %(T)s* Y = %(xb)d ? opA : opB;
""" % dict(tmp=total, ws=ws_gemm, nb=4 * max(1, len(blob.words)), T=T, swapped=int(pb.swapped),
           xb=int(bool(cand.swap))))
    if tA:
        src.append(_tplan_code("pA", tA))
        src.append("  %s* xt = reinterpret_cast<%s*>(tmp + %d);\n" % (T, T, offs["A"]))
        src.append("  if (int rc = tccg_transpose_%s(X, xt, 1, 0, &pA, s)) return lib_fail(rc, \"transposition of A\");\n" % fn)
        src.append("  X = xt;\n")
    if tB:
        src.append(_tplan_code("pB", tB))
        src.append("  %s* yt = reinterpret_cast<%s*>(tmp + %d);\n" % (T, T, offs["B"]))
        src.append("  if (int rc = tccg_transpose_%s(Y, yt, 1, 0, &pB, s)) return lib_fail(rc, \"transposition of B\");\n" % fn)
        src.append("  Y = yt;\n")
    src.append("""  tccg_gett_plan gp;
  memset(&gp, 0, sizeof(gp));
  gp.offAm = W(%(offAm)d); gp.offAk = W(%(offAk)d); gp.offBn = W(%(offBn)d); gp.offBk = W(%(offBk)d);
  gp.offCm = W(%(offCm)d); gp.offCn = W(%(offCn)d);
  gp.ordA = reinterpret_cast<const int32_t*>(W(%(ordA)d)); gp.ordB = reinterpret_cast<const int32_t*>(W(%(ordB)d));
  gp.Mp = %(Mp)d; gp.Np = %(Np)d; gp.Kp = %(Kp)d; gp.tile = %(tile)d; gp.modeA = %(modeA)d; gp.modeB = %(modeB)d;
  gp.groupM = %(groupM)d; gp.grid = %(grid)d; gp.walk = %(walk)d; gp.crun_m = %(crm)d; gp.crun_n = %(crn)d;
  gp.cvec = %(cvec)d; gp.ksplit = %(ksplit)d; gp.amax = nullptr;
  // the GEMM contraction's own A/B exchange
  const %(T)s* gA = %(gswap)d ? Y : X;
  const %(T)s* gB = %(gswap)d ? X : Y;
""" % dict(gt, Mp=g.Mp, Np=g.Np, Kp=g.Kp, tile=g.tile, modeA=g.modeA, modeB=g.modeB, groupM=g.groupM,
           grid=g.grid, walk=g.walk, crm=crun[0], crn=crun[1], cvec=g.cvec, ksplit=g.ksplit, T=T,
           gswap=int(gemm.problem.swapped)))
    if tC:
        src.append(_tplan_code("pC", tC))
        src.append("  %s* out = reinterpret_cast<%s*>(tmp + %d);\n" % (T, T, offs["C"]))
        src.append("  const %s alpha_ = alpha, beta_ = 0;\n" % T)
        src.append("  if (int rc = %s) return lib_fail(rc, \"GEMM\");\n" % gemm_call)
        src.append("  if (int rc = tccg_transpose_%s(out, C, 1, beta, &pC, s)) return lib_fail(rc, \"transposition of C\");\n" % fn)
    else:
        src.append("  %s* out = C;\n  const %s alpha_ = alpha, beta_ = beta;\n" % (T, T))
        src.append("  if (int rc = %s) return lib_fail(rc, \"GEMM\");\n" % gemm_call)
    src.append("  return 0;\n}\n")
    entry = ("int %s(const %s* A, const %s* B, %s* C, %s alpha, %s beta, %s* work_)" % (name, T, T, T, T, T, T))
    info = _info(contraction, prec, g.tile, g.BM, g.BN, g.BK, g.modeA, g.modeB)
    src.append(_tail(T, info, _desc(contraction_text), entry + " /* work_: unused, the library owns it */",
                     "(void)work_, gett_stream(A, B, C, alpha, beta, nullptr)"))
    return "".join(src)


def emit_log_source(contraction, cand, prec: int, contraction_text: str, tile: Optional[int] = None) -> str:
    """HIP source of the LoG method for `cand` (tccg_amd.loop_gemm): every
    GEMM instance a problem of one grouped launch (tccg_gett_group), the
    records built once per set of pointers, alpha and beta."""
    from .group import GettGroup, PACK_TF, PACK_TK
    from .loop_gemm import log_layout
    if prec not in _CTYPE:
        raise ValueError("LoG libraries are emitted for FP32 and FP64")
    T = _CTYPE[prec]
    esz = 8 if prec == F64 else 4
    pb = contraction.problem
    lay = log_layout(pb, cand, prec, tile)
    offAm, offAk, offBn, offBk, offCm, offCn = lay.tables
    g = GettGroup(prec, lay.tile)
    # pointers as byte offsets from the operand's (or C's) base
    for fa, fb, fc in lay.free_bases:
        a = g.operand(esz * fa, offAm, offAk, lay.kfastA, False, key=("A", fa))
        b = g.operand(esz * fb, offBn, offBk, lay.kfastB, True, key=("B", fb))
        g.problem(a, b, esz * fc, offCm, offCn, 1.0, 0.0, locality=fa)
    tables, ops, jobs, probs, work, need = g.records(0)
    nops, nprobs = len(g._ops), len(g._probs)
    name = cand.name
    src = ["// Generated by tccg_amd (gfx950 target): the LoG method, candidate %s; do not edit.\n" % name]
    for line in contraction_text.strip().splitlines():
        src.append("//   %s\n" % line)
    src.append("// %d GEMM instances of %d x %d, k = %d (contracted loops folded), packed tile %s\n"
               % (nprobs, int(np.count_nonzero(offAm != 0xffffffff)), int(np.count_nonzero(offBn != 0xffffffff)),
                  lay.K, "x".join(map(str, (g.BM, g.BN, g.BK)))))
    src.append(_PRELUDE)
    src.append(_words("log_tables", "int32_t", tables.tolist()))
    src.append(_words("log_jobs", "int32_t", jobs.reshape(-1).tolist()))
    src.append(_words("log_work", "int32_t", work.reshape(-1).tolist()))
    src.append("\nnamespace {\n")
    src.append(_host_array("kOpX", "int64_t", [ops[i].X or 0 for i in range(nops)]))
    src.append(_host_array("kOpF", "int64_t", [ops[i].offF or 0 for i in range(nops)]))
    src.append(_host_array("kOpK", "int64_t", [ops[i].offK or 0 for i in range(nops)]))
    src.append(_host_array("kOpOut", "int64_t", [ops[i].out for i in range(nops)]))
    src.append(_host_array("kOpMeta", "int32_t", [v for i in range(nops)
                                                  for v in (ops[i].fBlocks, ops[i].kTiles, ops[i].kfast, ops[i].isB)]))
    src.append(_host_array("kProbC", "int64_t", [probs[j].C or 0 for j in range(nprobs)]))
    src.append(_host_array("kProbCm", "int64_t", [probs[j].offCm or 0 for j in range(nprobs)]))
    src.append(_host_array("kProbCn", "int64_t", [probs[j].offCn or 0 for j in range(nprobs)]))
    src.append(_host_array("kProbPack", "int64_t", [v for j in range(nprobs) for v in (probs[j].packedA,
                                                                                        probs[j].packedB)]))
    src.append(_host_array("kProbMeta", "int32_t", [v for j in range(nprobs)
                                                    for v in (probs[j].tilesM, probs[j].tilesN, probs[j].kTiles)]))
    src.append(_STATE)
    src.append("""
constexpr int kOps = %(nops)d, kProbs = %(nprobs)d, kJobs = %(njobs)d, kWork = %(nwork)d;
constexpr size_t kWs = %(ws)d;

struct State {
  const unsigned char* tables;
  const tccg_group_job* jobs;
  const tccg_batch_work* work;
  tccg_group_operand* ops;      // device records for the pointers below
  tccg_group_problem* probs;
  void* ws;
  const void *A, *B, *C;        // pointers, alpha and beta the records hold
  double alpha, beta;
  bool filled;
  std::mutex mu;                // one record refresh at a time
  std::atomic<bool> ready;
};
State g_state[kMaxDevices];

int state_for_current_device(State** out) {
  int dev = 0;
  if (int rc = device_index(&dev)) return rc;
  State& t = g_state[dev];
  if (!t.ready.load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lock(g_init);
    if (!t.ready.load(std::memory_order_relaxed)) {
      const void* p = nullptr;
      if (int rc = symbol(p, HIP_SYMBOL(log_tables), %(nt)d)) return rc;
      t.tables = static_cast<const unsigned char*>(p);
      if (int rc = symbol(p, HIP_SYMBOL(log_jobs), %(njb)d)) return rc;
      t.jobs = static_cast<const tccg_group_job*>(p);
      if (int rc = symbol(p, HIP_SYMBOL(log_work), %(nwb)d)) return rc;
      t.work = static_cast<const tccg_batch_work*>(p);
      hipError_t e = hipMalloc(&t.ops, sizeof(tccg_group_operand) * kOps);
      if (e == hipSuccess) e = hipMalloc(&t.probs, sizeof(tccg_group_problem) * kProbs);
      if (e == hipSuccess) e = hipMalloc(&t.ws, kWs);
      if (e != hipSuccess) return fail((int)e, "hipMalloc of the LoG records or workspace failed");
      t.filled = false;
      t.ready.store(true, std::memory_order_release);
    }
  }
  *out = &t;
  return 0;
}

// Device records for these pointers, alpha and beta: the baked-in records
// hold byte offsets from the operand's (C's) base and from the tables.
int refresh(State* t, const void* opA, const void* opB, void* C, double alpha, double beta, hipStream_t s) {
  if (t->filled && t->A == opA && t->B == opB && t->C == C && t->alpha == alpha && t->beta == beta) return 0;
  std::vector<tccg_group_operand> ops(kOps);
  for (int i = 0; i < kOps; ++i) {
    tccg_group_operand& o = ops[i];
    const unsigned char* base = static_cast<const unsigned char*>(kOpMeta[4 * i + 3] ? opB : opA);
    o.X = base + kOpX[i];
    o.offF = reinterpret_cast<const uint32_t*>(t->tables + kOpF[i]);
    o.offK = reinterpret_cast<const uint32_t*>(t->tables + kOpK[i]);
    o.out = (uint64_t)kOpOut[i];
    o.fBlocks = kOpMeta[4 * i]; o.kTiles = kOpMeta[4 * i + 1]; o.kfast = kOpMeta[4 * i + 2]; o.isB = kOpMeta[4 * i + 3];
  }
  std::vector<tccg_group_problem> probs(kProbs);
  for (int j = 0; j < kProbs; ++j) {
    tccg_group_problem& q = probs[j];
    q.C = static_cast<unsigned char*>(C) + kProbC[j];
    q.offCm = reinterpret_cast<const uint32_t*>(t->tables + kProbCm[j]);
    q.offCn = reinterpret_cast<const uint32_t*>(t->tables + kProbCn[j]);
    q.packedA = (uint64_t)kProbPack[2 * j]; q.packedB = (uint64_t)kProbPack[2 * j + 1];
    q.tilesM = kProbMeta[3 * j]; q.tilesN = kProbMeta[3 * j + 1]; q.kTiles = kProbMeta[3 * j + 2]; q.pad = 0;
    q.alpha = alpha; q.beta = beta;
  }
  // earlier launches on this stream read the old records: let them finish,
  // then copy in stream order
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpyAsync(t->ops, ops.data(), sizeof(tccg_group_operand) * kOps, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(t->probs, probs.data(), sizeof(tccg_group_problem) * kProbs, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);   // the host vectors go out of scope
  if (e != hipSuccess) return fail((int)e, std::string("LoG record upload: ") + hipGetErrorString(e));
  t->A = opA; t->B = opB; t->C = C; t->alpha = alpha; t->beta = beta; t->filled = true;
  return 0;
}

}  // namespace

extern "C" {

int gett_stream(const %(T)s* A, const %(T)s* B, %(T)s* C, %(T)s alpha, %(T)s beta, void* stream) {
  if (!A || !B || !C) return fail(1, "tensor pointer is NULL");
  State* st = nullptr;
  if (int rc = state_for_current_device(&st)) return rc;
  hipStream_t s = (hipStream_t)stream;
  const %(T)s* opA = %(swapped)d ? B : A;
  const %(T)s* opB = %(swapped)d ? A : B;
  std::lock_guard<std::mutex> lock(st->mu);
  if (int rc = refresh(st, opA, opB, C, (double)alpha, (double)beta, s)) return rc;
  int rc = tccg_gett_group(%(prec)d, %(tile)d, st->ops, kOps, st->jobs, kJobs, st->probs, kProbs, st->work, kWork,
                           st->ws, kWs, s);
  if (rc) return lib_fail(rc, "grouped GETT");
  return 0;
}
""" % dict(nops=nops, nprobs=nprobs, njobs=int(jobs.shape[0]), nwork=int(work.shape[0]), ws=need,
           nt=4 * tables.size, njb=4 * max(1, jobs.size), nwb=4 * max(1, work.size), T=T,
           swapped=int(pb.swapped), prec=prec, tile=lay.tile))
    entry = "int %s(const %s* A, const %s* B, %s* C, %s alpha, %s beta)" % (name, T, T, T, T, T)
    BM, BN, BK = g.BM, g.BN, g.BK
    info = _info(contraction, prec, lay.tile, BM, BN, BK, lay.kfastA, lay.kfastB)
    src.append(_tail(T, info, _desc(contraction_text), entry, "gett_stream(A, B, C, alpha, beta, nullptr)"))
    return "".join(src)


def emit_header(T: str, method_entry: str) -> str:
    return """// Generated by tccg_amd (gfx950 target); do not edit.
// A, B, C are device pointers to dense column-major tensors (leftmost index
// stride-one); gett() runs asynchronously on the null stream.
#pragma once
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
%(entry)s;
void gett(%(T)s* A, %(T)s* B, %(T)s* C, const %(T)s alpha, const %(T)s beta);
int gett_stream(const %(T)s* A, const %(T)s* B, %(T)s* C, %(T)s alpha, %(T)s beta, void* stream);
int gett_info(int64_t* out, int n);  /* M, N, K, prec, tile, BM, BN, BK, modeA, modeB, |A|, |B|, |C| */
const char* gett_contraction(void);
const char* gett_last_error(void);
#ifdef __cplusplus
}
#endif
""" % dict(entry=method_entry, T=T)


def compile_source(src: Path, out: Path) -> Path:
    """hipcc into a shared library that calls libtccg_gfx950.so (found
    through the run path of the in-tree library)."""
    libdir = _build.LIB_DIR.resolve()
    cmd = [_build.HIPCC, *_build.HIP_FLAGS, "-I", str(_build.INCLUDE), "-shared", "-o", str(out), str(src),
           "-L", str(libdir), "-ltccg_gfx950", "-Wl,-rpath," + str(libdir)]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        msg = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
        raise _build.CompilationError("hipcc failed for %s:\n%s" % (src, msg))
    return out


def generate_method(kind: str, contraction, cand, prec: int, contraction_text: str, out_dir: Path,
                    gemm_plan=None, compile: bool = True) -> dict:
    """Write <base>.hip / <base>.hpp (and lib<base>.so) for the TTGT
    (kind "ttgt", base "ttgemmt") or LoG (kind "log", base "loopOverGemm")
    candidate, the reference's file names (tccg/tccg.py:171-185)."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    T = _CTYPE.get(prec)
    if T is None:
        raise ValueError("TTGT and LoG libraries are emitted for FP32 and FP64")
    if kind == "ttgt":
        if gemm_plan is None:
            from .contraction import Contraction
            sizes = {i.label: i.size for X in (contraction.problem.opA, contraction.problem.opB, contraction.problem.C)
                     for i in X.indices}
            gemm_plan = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes).plan(prec)
        text = emit_ttgt_source(contraction, cand, prec, gemm_plan, contraction_text)
        base = "ttgemmt"
        entry = "int ttgemmt_%s(const %s* A, const %s* B, %s* C, %s alpha, %s beta, %s* work_)" % (
            cand.name, T, T, T, T, T, T)
    elif kind == "log":
        text = emit_log_source(contraction, cand, prec, contraction_text)
        base = "loopOverGemm"
        entry = "int %s(const %s* A, const %s* B, %s* C, %s alpha, %s beta)" % (cand.name, T, T, T, T, T)
    else:
        raise ValueError("kind is 'ttgt' or 'log'")
    src = out_dir / (base + ".hip")
    hdr = out_dir / (base + ".hpp")
    src.write_text(text)
    hdr.write_text(emit_header(T, entry))
    res = {"source": src, "header": hdr, "library": None, "entry": entry}
    if compile:
        res["library"] = compile_source(src, out_dir / ("lib%s.so" % base))
    return res


# Libraries built with the package (used by the GPU tests of the emitted
# methods): (directory, workload text builder, method, precision).
def method_library_specs():
    from .workloads import scaled
    small = scaled("abcd-aebf-dfce", 0.34)        # 24/16
    swapped = "C[a,b,c] = 1.5 * B[k,a,b] * A[c,k] + 0.5 * C[a,b,c]\na = 24\nb = 8\nc = 40\nk = 12\n"
    return [("ttgt_f32", small, "ttgt", F32), ("ttgt_f64", small, "ttgt", F64),
            ("log_f32", small, "log", F32), ("log_f64", small, "log", F64),
            ("ttgt_swapped_f32", swapped, "ttgt", F32), ("log_swapped_f32", swapped, "log", F32)]


def build_method_libraries(out_root: Optional[Path] = None) -> List[Path]:
    """Emit and compile the TTGT and LoG libraries of method_library_specs()
    (the model's best candidate of each) under tccg_amd/_generated/methods."""
    from concurrent.futures import ThreadPoolExecutor
    from .codegen import GENERATED_DIR
    from .contraction import Contraction
    from .loop_gemm import best_log_candidates, log_candidates
    from .ttgt import best_candidates, ttgt_candidates
    root = Path(out_root) if out_root is not None else GENERATED_DIR / "methods"
    jobs = []
    for dname, text, kind, prec in method_library_specs():
        c = Contraction.from_tccg(text)
        if kind == "ttgt":
            cand = best_candidates(ttgt_candidates(c.problem, 8 if prec == F64 else 4), 1)[0]
        else:
            cand = best_log_candidates(log_candidates(c.problem), 1)[0]
        jobs.append((kind, c, cand, prec, text, root / dname))

    def run(job):
        kind, c, cand, prec, text, d = job
        return generate_method(kind, c, cand, prec, text, d)["library"]

    with ThreadPoolExecutor(max_workers=min(6, len(jobs))) as ex:
        return list(ex.map(run, jobs))
