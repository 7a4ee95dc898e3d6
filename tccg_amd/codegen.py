"""gfx950 code generator: one HIP source per contraction, keeping the
reference's generated entry point

    void gett(T* A, T* B, T* C, const T alpha, const T beta)

(tccg/gett.py:342-344, emitted at tccg/gett.py:1240-1243) with every extent
fixed at generation time, as the reference does (tccg/gett.py:328-330).  The
emitted source instantiates the hand-written kernel of csrc/gett_core.hpp for
the chosen plan (tile shape and operand load modes become template
arguments) and embeds the plan's offset tables as device constant data; the
library is compiled with hipcc into <name>.so next to the source, as the
reference compiles its gett<N>.cpp with make (tccg/tccg.py:244-276).

Besides gett(), the generated library exports
    int gett_stream(const T* A, const T* B, T* C, T alpha, T beta, void* stream)
    int gett_info(int64_t* out, int n)   // M, N, K, prec, tile, BM, BN, BK, modeA, modeB,
                                         // elements of A, B, C (the .tccg's tensors)
    const char* gett_contraction(void)
    const char* gett_last_error(void)
and the header <name>.hpp declares them.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
from pathlib import Path
from typing import Optional

import numpy as np

from . import build as _build
from .plan import BF16, BF16X3, F32, F64, FP8, PREC_NAMES, WAVES, GettPlan

_CTYPE = {F32: "float", F64: "double", BF16X3: "float", FP8: "float", BF16: "float"}
_PREC_ENUM = {F32: "tccg::F32", F64: "tccg::F64", BF16X3: "tccg::BF16X3", FP8: "tccg::FP8", BF16: "tccg::BF16"}

GENERATED_DIR = Path(__file__).resolve().parent / "_generated"


def _array(name: str, ctype: str, values: np.ndarray) -> str:
    """A device table with external linkage, kept by the device compile: a table
    that device code never names directly would otherwise be dropped from the
    code object while the host still registers its symbol, and a namespace-scope
    const would get internal linkage (a local symbol the loader cannot look up
    by name).  The kernels only read it."""
    vals = ",".join(str(int(v)) for v in values.tolist())
    # 16-byte aligned: the store-wave kernel copies C's offsets in 16-byte pieces
    return "__device__ __attribute__((used, aligned(16))) %s %s[%d] = {%s};\n" % (ctype, name, values.size, vals)


def _table_names(name: str):
    base = "tccg_" + "".join(ch if ch.isalnum() else "_" for ch in name)
    return {k: "%s_%s" % (base, k) for k in ("offAm", "offAk", "offBn", "offBk", "offCm", "offCn", "ordA", "ordB")}


def _part_offset(plan: GettPlan) -> int:
    """Workspace offset of the partial tiles (split-k and stream-k plans):
    after the packed operands, at a 256-byte boundary
    (tccg_packed_workspace_bytes)."""
    per = {F32: 4, F64: 8, BF16X3: 6, BF16: 2, FP8: 1}[plan.prec] * plan.Kp
    ab = -(-(per * plan.Mp) // 256) * 256 + per * plan.Np
    return -(-ab // 256) * 256


def _part_expr(plan: GettPlan, T: str) -> str:
    """The partial-tile buffer of a split-k or stream-k plan."""
    if plan.ksplit <= 1 and plan.walk != 5:
        return "nullptr"
    return "reinterpret_cast<%s*>(reinterpret_cast<unsigned char*>(t->ws) + %d)" % (T, _part_offset(plan))


def _flags_offset(plan: GettPlan) -> int:
    """Workspace offset of a stream-k plan's flag words: after one partial
    tile per workgroup, at a 256-byte boundary."""
    esz = 8 if plan.prec == F64 else 4
    return _part_offset(plan) + -(-(plan.grid * plan.BM * plan.BN * esz) // 256) * 256


def _reduce_launch(plan: GettPlan, T: str) -> str:
    if plan.ksplit <= 1:
        return ""
    groups = plan.ntiles() * plan.BN * (plan.BM // 4)
    grid = min(-(-groups // 256), 2048)
    return ("""  hipLaunchKernelGGL((tccg::split_reduce_kernel<%s, %d, %d>), dim3(%d), dim3(256), 0, s, p);
  e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("split reduction launch: ") + hipGetErrorString(e));
""" % (T, plan.BM, plan.BN, grid))


def emit_source(plan: GettPlan, contraction_text: str, name: str = "gett") -> str:
    """HIP source of the per-contraction library for `plan`."""
    pb = plan.problem
    T = _CTYPE[plan.prec]
    WM, WN = WAVES.get((plan.prec, plan.tile), (2, 2))
    walk = 2 if plan.walk == 2 else (1 if plan.grid else 0)
    vector = plan.walk == 3
    if vector:
        kern = "tccg::gett_vector_kernel<%s, %d, %d, %d, %d>" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, 4 if plan.modeA == 0 else 1)
        shape = "void"   # no tile shape: the kernel's LDS is static
    elif plan.packed and plan.walk == 4:
        WM, WN = 2, 2
        from .plan import sw_stages
        stages = sw_stages(plan.prec, plan.BM, plan.BN, plan.BK)
        kern = "tccg::gett_packed_sw_kernel<%s, %d, %d, %d, %d>" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, stages)
        shape = "tccg::SwShape<%s, %d, %d, %d, %d>" % (_PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, stages)
    elif plan.packed and plan.walk == 5:
        WM, WN = 2, 2
        kern = "tccg::gett_packed_sk_kernel<%s, %d, %d, %d, 2, 2, 3>" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK)
        shape = "tccg::PackedShape<%s, %d, %d, %d, 3>" % (_PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK)
    elif plan.packed:
        WM, WN = 2, 2
        kern = "tccg::gett_packed_kernel<%s, %d, %d, %d, 2, 2, 3, %s>" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, "true" if plan.grid else "false")
        shape = "tccg::PackedShape<%s, %d, %d, %d, 3>" % (_PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK)
    else:
        kern = "tccg::gett_kernel<%s, %d, %d, %d, %d, %d, %d, %d, %d>" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, WM, WN, plan.modeA, plan.modeB, walk)
        shape = "tccg::TileShape<%s, %d, %d, %d, %d, %d>" % (_PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, WM, WN)
    desc = contraction_text.strip().replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n")
    swap = "1" if pb.swapped else "0"
    out = []
    out.append("// Generated by tccg_amd (gfx950 GETT target); do not edit.\n")
    out.append("// Contraction:\n")
    for line in contraction_text.strip().splitlines():
        out.append("//   %s\n" % line)
    out.append("// Plan: %s\n" % plan.describe())
    out.append('#include <hip/hip_runtime.h>\n#include <stdint.h>\n\n#include <atomic>\n#include <mutex>\n#include <string>\n#include "%s"\n\n'
               % ("gett_vector.hpp" if vector else "gett_packed_sw.hpp\"\n#include \"gett_pack.hpp" if plan.packed
                  else "gett_core.hpp"))
    tn = _table_names(name)
    tables = (("offAm", plan.offAm, "uint32_t"), ("offAk", plan.offAk, "uint32_t"),
              ("offBn", plan.offBn, "uint32_t"), ("offBk", plan.offBk, "uint32_t"),
              ("offCm", plan.offCm, "uint32_t"), ("offCn", plan.offCn, "uint32_t"),
              ("ordA", plan.ordA, "int"), ("ordB", plan.ordB, "int"))
    for key, arr, ct in tables:
        out.append(_array(tn[key], ct, arr))
    out.append("\nnamespace {\n")
    out.append("""
constexpr int kMaxDevices = 64;
struct DeviceTables {
  const uint32_t *Am, *Ak, *Bn, *Bk, *Cm, *Cn;
  const int *oA, *oB;
  float* amax;
  void* ws;
  std::atomic<bool> ready;   // set once the fields above hold this device's addresses
};
DeviceTables g_tables[kMaxDevices];
std::mutex g_init;           // serialises first-call set-up (allocations included)
thread_local std::string g_err;

int fail(int code, const std::string& m) { g_err = m; return code ? code : 1; }

// Device address of a table, after checking that the code object holds it
// with the expected size.
template <class P>
int symbol(P& dst, const void* sym, size_t bytes) {
  void* p = nullptr;
  size_t n = 0;
  hipError_t e = hipGetSymbolSize(&n, sym);
  if (e != hipSuccess) return fail((int)e, std::string("hipGetSymbolSize: ") + hipGetErrorString(e));
  if (n != bytes) return fail(1, "device table has an unexpected size");
  e = hipGetSymbolAddress(&p, sym);
  if (e != hipSuccess) return fail((int)e, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  if (!p) return fail(1, "device table address is NULL");
  dst = reinterpret_cast<P>(p);
  return 0;
}

int tables_for_current_device(DeviceTables** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail((int)e, "hipGetDevice failed");
  if (dev < 0 || dev >= kMaxDevices) return fail(1, "device index out of range");
  DeviceTables& t = g_tables[dev];
  if (!t.ready.load(std::memory_order_acquire)) {
    std::lock_guard<std::mutex> lock(g_init);
    if (t.ready.load(std::memory_order_relaxed)) {
      *out = &t;
      return 0;
    }
    int rc = %s;
    if (rc) return rc;
""" % " ||\n             ".join("symbol(t.%s, HIP_SYMBOL(%s), %d)" % (fld, tn[key], arr.size * 4)
                                  for fld, (key, arr, ct) in zip(("Am", "Ak", "Bn", "Bk", "Cm", "Cn", "oA", "oB"), tables)))
    if plan.prec == FP8:
        out.append("""    e = hipMalloc(&t.amax, 2 * sizeof(float));
    if (e != hipSuccess) return fail((int)e, "hipMalloc of the FP8 amax pair failed");
""")
    if plan.packed:
        out.append("""    e = hipMalloc(&t.ws, %d);
    if (e != hipSuccess) return fail((int)e, "hipMalloc of the packed-operand workspace failed");
""" % plan.workspace_bytes())
    out.append("""    t.ready.store(true, std::memory_order_release);
  }
  *out = &t;
  return 0;
}
""")
    if plan.prec == FP8:
        out.append("""
__global__ void amax_zero(unsigned* o) { if (threadIdx.x == 0) { o[0] = 0u; o[1] = 0u; } }
__global__ void amax_reduce(const float* __restrict__ x, int64_t n, unsigned* o) {
  __shared__ float red[4];
  float m = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(x[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    atomicMax(o, __float_as_uint(b));
  }
}
""")
    out.append("}  // namespace\n\n")
    out.append('extern "C" {\n\n')
    out.append("int gett_stream(const {T}* A, const {T}* B, {T}* C, {T} alpha, {T} beta, void* stream) {{\n".format(T=T))
    if vector:
        lds = "0"
    elif plan.packed:
        lds = "S::LDS_BYTES"
    elif walk == 2:
        lds = "(int)tccg::stream_lds_bytes<%s, %d, %d, %d, %d, %d>(%d)" % (
            _PREC_ENUM[plan.prec], plan.BM, plan.BN, plan.BK, WM, WN, plan.Np // plan.BN)
    elif plan.Kp == plan.BK:
        lds = "S::LDS_BYTES / 2"   # a single k-step never touches the second stage
    else:
        lds = "S::LDS_BYTES"
    out.append(("""  using S = %s;
  constexpr int kLds = """ + lds + """;
  if (!A || !B || !C) return fail(1, "tensor pointer is NULL");
  DeviceTables* t = nullptr;
  int rc = tables_for_current_device(&t);
  if (rc) return rc;
  auto kern = %s;
  // the dynamic-LDS limit is a per-device attribute: one bit per device id
  static std::atomic<uint64_t> lds_set{0};
  int lds_dev = 0;
  if (hipGetDevice(&lds_dev) != hipSuccess) return fail(1, "hipGetDevice failed");
  const uint64_t dev_bit = lds_dev < 64 ? (uint64_t{1} << lds_dev) : 0;
  if (!dev_bit || !(lds_set.load(std::memory_order_acquire) & dev_bit)) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, kLds);
    if (e != hipSuccess) return fail((int)e, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    lds_set.fetch_or(dev_bit, std::memory_order_release);
  }
  const %s* opA = %s ? B : A;
  const %s* opB = %s ? A : B;
""") % (shape, kern, T, swap, T, swap))
    if not plan.packed and plan.modeA in (0, 2):
        out.append('  if (reinterpret_cast<uintptr_t>(opA) & 15u) return fail(1, "operand A of the GETT must be 16-byte aligned");\n')
    if not plan.packed and plan.modeB in (0, 2):
        out.append('  if (reinterpret_cast<uintptr_t>(opB) & 15u) return fail(1, "operand B of the GETT must be 16-byte aligned");\n')
    if (vector and plan.modeA == 0) or (plan.packed and (plan.cvec or plan.crun[0])):
        out.append('  if (reinterpret_cast<uintptr_t>(C) & 15u) return fail(1, "C must be 16-byte aligned");\n')
    out.append("  hipStream_t s = (hipStream_t)stream;\n")
    if plan.prec == FP8:
        out.append("""  hipLaunchKernelGGL(amax_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(t->amax));
  hipLaunchKernelGGL(amax_reduce, dim3(256), dim3(256), 0, s, opA, (int64_t)%d, reinterpret_cast<unsigned*>(t->amax));
  hipLaunchKernelGGL(amax_reduce, dim3(256), dim3(256), 0, s, opB, (int64_t)%d, reinterpret_cast<unsigned*>(t->amax) + 1);
""" % (pb.opA.size(), pb.opB.size()))
    if plan.packed:
        TL = {F32: "float", F64: "double", BF16X3: "__bf16", BF16: "__bf16", FP8: "uint8_t"}[plan.prec]
        esz = {F32: 4, F64: 8, BF16X3: 2, BF16: 2, FP8: 1}[plan.prec]
        planes = 3 if plan.prec == BF16X3 else 1
        offB = -(-(planes * plan.Mp * plan.Kp * esz) // 256) * 256
        kfA = 1 if plan.modeA in (2, 3) else 0
        kfB = 1 if plan.modeB in (2, 3) else 0
        amax = "t->amax" if plan.prec == FP8 else "nullptr"
        flags = ("reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(t->ws) + %d)" % _flags_offset(plan)
                 if plan.walk == 5 else "nullptr")
        out.append("""  %(TL)s* wA = reinterpret_cast<%(TL)s*>(t->ws);
  %(TL)s* wB = reinterpret_cast<%(TL)s*>(reinterpret_cast<unsigned char*>(t->ws) + %(offB)d);
  // both operands in one launch, which also clears the stream-k flags
  tccg::PackPair<%(T)s, %(TL)s> q;
  q.X[0] = opA; q.X[1] = opB; q.offF[0] = t->Am; q.offF[1] = t->Bn; q.offK[0] = t->Ak; q.offK[1] = t->Bk;
  q.amax[0] = %(amax)s; q.amax[1] = %(amaxB)s; q.out[0] = wA; q.out[1] = wB;
  q.kfast[0] = %(kfA)d; q.kfast[1] = %(kfB)d; q.nF[0] = %(gA)d; q.nF[1] = %(gB)d; q.nK = %(gK)d; q.kTiles = %(KT)d;
  q.clear = %(flags)s; q.nclear = %(nclear)d;
  hipLaunchKernelGGL((tccg::pack_pair_kernel<%(P)s, %(BM)d, %(BN)d, %(BK)d>), dim3(%(blocks)d), dim3(256), 0, s, q);
  opA = reinterpret_cast<const %(T)s*>(wA);
  opB = reinterpret_cast<const %(T)s*>(wB);
""" % dict(TL=TL, offB=offB, P=_PREC_ENUM[plan.prec], BM=plan.BM, BN=plan.BN, BK=plan.BK, gA=plan.Mp // 64,
           gB=plan.Np // 64, gK=plan.Kp // 32, KT=plan.Kp // plan.BK, kfA=kfA, kfB=kfB, amax=amax,
           amaxB=(amax + " + 1") if plan.prec == FP8 else "nullptr", T=T, flags=flags,
           nclear=plan.grid + 1 if plan.walk == 5 else 0, blocks=(plan.Mp // 64 + plan.Np // 64) * (plan.Kp // 32)))
    out.append("""  tccg::GettParams<%s> p;
  p.A = opA; p.B = opB; p.C = C;
  p.offAm = t->Am; p.offAk = t->Ak; p.offBn = t->Bn; p.offBk = t->Bk; p.offCm = t->Cm; p.offCn = t->Cn;
  p.ordA = t->oA; p.ordB = t->oB;
  p.tilesM = %d; p.tilesN = %d; p.kTiles = %d; p.groupM = %d;
  p.alpha = alpha; p.beta = beta;
  p.amax = %s;
  p.crunM = %d; p.crunN = %d; p.cvec = %d;
  p.ksplit = %d; p.part = %s; p.flags = %s;
  hipLaunchKernelGGL(kern, dim3(%d), dim3(%d), kLds, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
%s  return 0;
}

""" % (T, plan.Mp // plan.BM, plan.Np // plan.BN, plan.Kp // plan.BK, plan.groupM,
       "t->amax" if plan.prec == FP8 else "nullptr", plan.crun[0], plan.crun[1], plan.cvec,
       plan.ksplit, _part_expr(plan, T),
       ("reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(t->ws) + %d)" % _flags_offset(plan)
        if plan.walk == 5 else "nullptr"),
       plan.grid if plan.grid else plan.ntiles() * plan.ksplit, 512 if plan.walk == 4 else 256,
       _reduce_launch(plan, T)))
    out.append("void gett({T}* A, {T}* B, {T}* C, const {T} alpha, const {T} beta) {{\n"
               "  (void)gett_stream(A, B, C, alpha, beta, nullptr);\n}}\n\n".format(T=T))
    # sizes of the .tccg's A, B and C (gett()'s arguments in their positional order)
    sizeA, sizeB = (pb.opB.size(), pb.opA.size()) if pb.swapped else (pb.opA.size(), pb.opB.size())
    info = [pb.M, pb.N, pb.K, plan.prec, plan.tile, plan.BM, plan.BN, plan.BK, plan.modeA, plan.modeB,
            sizeA, sizeB, pb.C.size()]
    out.append("int gett_info(int64_t* out, int n) {\n  const int64_t v[%d] = {%s};\n"
               "  for (int i = 0; i < n && i < %d; ++i) out[i] = v[i];\n  return %d;\n}\n\n"
               % (len(info), ",".join(str(int(x)) for x in info), len(info), len(info)))
    out.append('const char* gett_contraction(void) { return "%s"; }\n\n' % desc)
    out.append("const char* gett_last_error(void) { return g_err.c_str(); }\n\n")
    out.append('}  // extern "C"\n')
    return "".join(out)


def emit_header(plan: GettPlan, name: str = "gett") -> str:
    T = _CTYPE[plan.prec]
    return """// Generated by tccg_amd (gfx950 GETT target); do not edit.
// A, B, C are device pointers to dense column-major tensors (leftmost index
// stride-one); gett() runs asynchronously on the null stream.
#pragma once
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
void gett(%(T)s* A, %(T)s* B, %(T)s* C, const %(T)s alpha, const %(T)s beta);
int gett_stream(const %(T)s* A, const %(T)s* B, %(T)s* C, %(T)s alpha, %(T)s beta, void* stream);
int gett_info(int64_t* out, int n);  /* M, N, K, prec, tile, BM, BN, BK, modeA, modeB, |A|, |B|, |C| */
const char* gett_contraction(void);
const char* gett_last_error(void);
#ifdef __cplusplus
}
#endif
""" % {"T": T}


def compile_source(src: Path, out: Path) -> Path:
    cmd = [_build.HIPCC, *_build.HIP_FLAGS, "-I", str(_build.CSRC), "-shared", "-o", str(out), str(src)]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        msg = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
        raise _build.CompilationError("hipcc failed for %s:\n%s" % (src, msg))
    return out


def generate(plan: GettPlan, contraction_text: str, out_dir: Path, name: str = "gett",
             compile: bool = True) -> dict:
    """Write <name>.hip and <name>.hpp into out_dir and compile lib<name>.so."""
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    src = out_dir / (name + ".hip")
    hdr = out_dir / (name + ".hpp")
    src.write_text(emit_source(plan, contraction_text, name))
    hdr.write_text(emit_header(plan, name))
    res = {"source": src, "header": hdr, "library": None}
    if compile:
        res["library"] = compile_source(src, out_dir / ("lib%s.so" % name))
    return res


def workload_library_name(workload: str, prec: int) -> str:
    return "gett_%s_%s" % (workload.replace("-", "_"), PREC_NAMES[prec])


def packed_suffix(packed) -> str:
    """Library-name suffix of a packed plan's options (none for planner plans)."""
    if packed is None:
        return ""
    if packed.get("store_waves"):
        return "_storewaves"
    if packed.get("ksplit", 1) > 1:
        return "_ksplit%d" % packed["ksplit"]
    if packed.get("stream_k"):
        return "_streamk%d" % packed["stream_k"]
    return "_packed"


# Contractions built into libraries besides the named workloads: the first
# adjoint of abcd-aklb-lckd (dA[a,k,l,b] = dC[a,b,c,d] * B[l,c,k,d]) at
# extents 24/16, M*N = 576*256 with k = 576.
SPLIT_K_TEXTS = {
    "aklb-abcd-lckd": "C[a,k,l,b] = A[a,b,c,d] * B[l,c,k,d]\na = 24\nb = 24\nc = 24\nd = 24\nk = 16\nl = 16\n",
}
# A contraction whose A has no free index (C's indices all come from B, the
# n set is empty): the generated entry point exchanges the operands.
VECTOR_A_TEXTS = {
    "ji-kh-hikj": "C[j,i] = 1.5 * A[k,h] * B[h,i,k,j]\nj = 37\ni = 5\nk = 21\nh = 3\n",
}
# Index sets at their edges, each compiled with the planner's best plan: an
# empty n set (a vector kernel), an extent-1 index among shuffled orders, A
# without a free index (packed store waves, operands exchanged), fused
# indices with exchanged operands, an empty k set (outer product).  (A
# tensor with one index needs a label of two characters: the .tccg tensor
# pattern, tccg/tccg.py:574, matches at least two.)
EDGE_TEXTS = {
    "heg-ehakg-ak": "C[h,e,g] = A[e,h,ak,g] * B[ak]\nh = 30\ne = 7\ng = 36\nak = 8\n",
    "hge-ihed-dig": "C[h,g,e] = A[i,h,e,d] * B[d,i,g]\nh = 20\ng = 4\ne = 21\ni = 6\nd = 1\n",
    "cg-jk-gjkc": "C[c,g] = A[jk] * B[g,jk,c]\nc = 42\ng = 11\njk = 26\n",
    "jch-dchf-djf": "C[j,c,h] = A[d,c,h,f] * B[d,j,f]\nj = 32\nc = 25\nh = 24\nd = 7\nf = 25\n",
    "icfe-fe-ci": "C[i,c,f,e] = A[f,e] * B[c,i]\ni = 42\nc = 5\nf = 43\ne = 23\n",
}


def build_workload_libraries(force: bool = False, workloads=None, precs=(F32,)) -> list:
    """Generate and compile the per-contraction libraries of the named
    workloads (used by the GPU tests of the generated entry point)."""
    from concurrent.futures import ThreadPoolExecutor
    from .contraction import Contraction
    from .workloads import WORKLOADS, scaled
    from .plan import build_packed_plan
    jobs = []
    names = workloads or ["example", "cpu48", "abcd-aebf-dfce", "abc-akb-kc"]
    todo = [(wl, prec, None) for wl in names for prec in precs]
    texts = dict(WORKLOADS)
    if workloads is None:
        # packed-operand libraries: split-BF16 on the compute-bound contraction,
        # FP8 (persistent walk) on the precision contraction, FP8 store waves,
        # and split-k and stream-k plans on the first adjoint of abcd-aklb-lckd
        # (small M*N, long k)
        texts.update(SPLIT_K_TEXTS)
        texts.update(VECTOR_A_TEXTS)
        texts.update(EDGE_TEXTS)
        todo += [(wl, F32, None) for wl in list(VECTOR_A_TEXTS) + list(EDGE_TEXTS)]
        todo += [("abcd-aebf-dfce", BF16X3, dict(tile=0, persistent=False)),
                 ("abcdef-dgfa-gceb", FP8, dict(tile=0, persistent=True)),
                 ("abcdef-dgfa-gceb", FP8, dict(tile=4, store_waves=True)),
                 ("aklb-abcd-lckd", F32, dict(tile=0, ksplit=4)),
                 ("aklb-abcd-lckd", F32, dict(tile=2, stream_k=1))]
    for wl, prec, packed in todo:
        text = texts[wl]
        c = Contraction.from_tccg(text)
        plan = c.plan(prec) if packed is None else build_packed_plan(c.problem, prec, **packed)
        nm = workload_library_name(wl, prec) + packed_suffix(packed)
        src_text = emit_source(plan, text, nm)
        digest = hashlib.sha256((src_text + "".join((_build.CSRC / h).read_text() for h in _build.LIB_HEADERS)).encode()).hexdigest()[:16]
        lib = GENERATED_DIR / ("lib%s.so" % nm)
        stamp = GENERATED_DIR / ("lib%s.digest" % nm)
        if not force and lib.exists() and stamp.exists() and stamp.read_text().strip() == digest:
            continue
        jobs.append((plan, text, nm, digest))

    def run(job):
        plan, text, nm, digest = job
        generate(plan, text, GENERATED_DIR, nm)
        (GENERATED_DIR / ("lib%s.digest" % nm)).write_text(digest + "\n")
        return nm

    GENERATED_DIR.mkdir(parents=True, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, max(1, len(jobs)))) as ex:
        return list(ex.map(run, jobs))
