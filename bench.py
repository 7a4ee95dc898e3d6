#!/usr/bin/env python3
"""Benchmark of the gfx950 GETT path on one MI355X.

A step is one GETT launch over one contraction whose operands are already
resident in HBM.  The headline workload is the compute-bound plain-FP32
contraction C[a,b,c,d] = A[a,e,b,f] * B[d,f,c,e] (a=b=c=d=72, e=f=48,
alpha=1, beta=0, inputs iid uniform[-1,1]); the same JSON line carries the
bandwidth-bound contraction C[a,b,c] = A[a,k,b] * B[k,c] (384/8), the FP64
variant, the split-BF16 and FP8 forms with their error against FP64, the
roofline of the dominant kernel and the CPU baseline.

    python bench.py [--gpus 1] [--steps 20] [--warmup 5]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GFLOP/s + roofline %; split-BF16 vs plain FP32 GETT at equal error vs FP64"
PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6, "bf16x3": 2516.6 / 6.0, "bf16": 2516.6, "fp8": 5033.0}
PEAK_HBM_GBS = 8000.0


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--max-implementations", type=int, default=16,
                    help="candidates timed by the autotuner per workload (tccg --maxImplementations)")
    ap.add_argument("--only", default=None,
                    help="run one form only: f32, bw, f64, bf16x3, bf16, fp8, fp8_f32, batch, grad, ttgt, log, mfma (profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-seconds", type=float, default=15.0)
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (PMC calibration)")
    ap.add_argument("--candidate", type=int, default=None,
                    help="skip the autotuner: run the K-th model-ranked candidate plan (profiling a chosen plan)")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def rel_err(x, ref) -> float:
    import torch
    d = (x.double() - ref).norm()
    return float(d / ref.norm())


def _einsum_eq(c) -> str:
    import string
    letters = {}
    pool = iter(string.ascii_letters)
    for lab in [i.label for i in c.C0.indices + c.A0.indices + c.B0.indices]:
        letters.setdefault(lab, next(pool))
    rev = lambda T: "".join(letters[i.label] for i in reversed(T.indices))
    return "%s,%s->%s" % (rev(c.A0), rev(c.B0), rev(c.C0))


def fp64_reference(c, A, B):
    """alpha*A.B in float64 on the GPU with torch.einsum over contiguous
    operands (an independent library path, not this backend)."""
    import torch
    return torch.einsum(_einsum_eq(c), A.double().contiguous(), B.double().contiguous()) * c.alpha


def seeded_inputs(c, dt, dist: str = "uniform", seed: int = 0):
    """A then B drawn with numpy.random.default_rng(seed) in FP64 (iid
    uniform[-1, 1] or N(0, 1)), cast to dt, uploaded to the GPU."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    sa, sb, _ = c.shapes()
    draw = (lambda s: rng.uniform(-1.0, 1.0, s)) if dist == "uniform" else (lambda s: rng.standard_normal(s))
    A = torch.from_numpy(draw(sa)).to(dt).cuda()
    B = torch.from_numpy(draw(sb)).to(dt).cuda()
    return A, B


SETTLE_SECONDS = 0.05


def settle(fn, warmup: int) -> int:
    """The untimed warmup: `warmup` calls, then more until the GPU has run
    calls for SETTLE_SECONDS, so that its clocks have left the idle state
    the host-side tuning left them in (measured on MI355X: the first 20 ms
    of calls after idle run up to 13 % slower).  Returns the calls made."""
    import torch
    n = 0
    t0 = time.perf_counter()
    while n < warmup or time.perf_counter() - t0 < SETTLE_SECONDS:
        fn()
        n += 1
        if n >= warmup:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    return n


def refine(ms, A, B, C, steps: int):
    """The fastest of the autotuner's best few plans by graph-replayed calls
    (pack kernels + GETT, `steps` per replay, the minimum of three, each
    replay checked against an eager call): the eager timings of the
    autotuner tell plans within a percent or two apart less reliably."""
    from tccg_amd.runtime import DevicePlan
    best, best_t = ms[0].plan, float("inf")
    for m in ms:
        dp = DevicePlan(m.plan)
        settle(lambda: dp.run(A, B, C), 2)
        t = graph_call_time(lambda: dp.run(A, B, C), C, steps)
        if t is not None and t < best_t:
            best, best_t = m.plan, t
        del dp
    return best


def timed_steps(dp, A, B, C, steps: int, warmup: int):
    """(wall seconds of `steps` back-to-back calls bracketed by synchronize,
    average per-call GPU time in seconds from HIP events recorded on the
    launch stream before the first and after the last call: the GPU stays
    busy, so host launch latency is not counted, and the warmup calls made)."""
    import torch
    s = torch.cuda.current_stream()
    nwarm = settle(lambda: dp.run(A, B, C), warmup)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(steps):
        dp.run(A, B, C)
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / steps * 1e-3, nwarm


def graph_call_time(fn, C, steps: int, reps: int = 3):
    """GPU time (seconds) of one call of fn(): `steps` calls captured in a HIP
    graph and replayed, so the launches run back to back with no host work
    between them (the eager loop leaves gaps before kernels of a few tens of
    microseconds).  Every replay must leave C bit-identical to an eager call
    (the calls are idempotent: beta = 0).  None if capture fails."""
    import torch
    fn()
    torch.cuda.synchronize()
    ref = C.clone()
    try:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                for _ in range(steps):
                    fn()
        torch.cuda.current_stream().wait_stream(s)
    except Exception:
        torch.cuda.synchronize()
        return None
    best = float("inf")
    for _ in range(reps):
        C.zero_()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        g.replay()
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1) * 1e-3 / steps)
        if not torch.equal(C, ref):
            raise RuntimeError("graph replay differs from the eager result")
    del g
    return best


def kernel_only(dp, C, steps: int):
    """Average HIP-event duration (seconds) of the GETT kernel alone: for a
    packed plan the kernel over the operands the last run() packed, for an
    unpacked plan the single launch of run() itself (None then)."""
    import torch
    if not dp.plan.packed:
        return None
    s = torch.cuda.current_stream()
    dp.run_prepacked(C)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(steps):
        dp.run_prepacked(C)
    e1.record(s)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps * 1e-3


def cold_times(dp, A, B, C, reps: int = 10):
    """(minimum, median) call time after writing a 512 MiB buffer before each
    call (evicts L2 and the 256 MiB Infinity Cache), the reference's
    trash-then-time method (tccg/tccg.py:725-737, tccg/tccg.py:1032-1058)."""
    import statistics
    import torch
    trash = torch.empty(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream()
    times = []
    for r in range(reps):
        trash.fill_(float(r))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        dp.run(A, B, C)
        e1.record(s)
        e1.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    del trash
    return min(times), statistics.median(times)


def run_form(name, wl, prec, args, log):
    import torch
    from tccg_amd import Contraction
    from tccg_amd.autotune import autotune
    from tccg_amd.runtime import DevicePlan
    from tccg_amd.workloads import get

    c = Contraction.from_tccg(get(wl))
    dt = torch.float64 if prec == "f64" else torch.float32
    A, B = seeded_inputs(c, dt, "normal" if wl == "abcdef-dgfa-gceb" else "uniform")
    C = torch.zeros(c.shapes()[2], dtype=dt, device="cuda")
    from tccg_amd.plan import PREC_BY_NAME
    p = PREC_BY_NAME[prec]
    t0 = time.perf_counter()
    if args.candidate is not None:
        from tccg_amd.plan import candidate_plans
        best = candidate_plans(c.problem, p)[args.candidate]
        ms = [best]
    else:
        ms = autotune(c, p, max_implementations=args.max_implementations, reps=3, inputs=(A, B, C),
                      verbose=args.verbose)
        best = refine(ms[:3], A, B, C, args.steps)
    tune_s = time.perf_counter() - t0
    dp = DevicePlan(best)
    wall, call, nwarm = timed_steps(dp, A, B, C, args.steps, args.warmup)
    # GPU time per call and per GETT kernel without host gaps between launches
    g_call = graph_call_time(lambda: dp.run(A, B, C), C, args.steps)
    if g_call is not None:
        call = g_call
    kern = None
    if dp.plan.packed:
        kern = graph_call_time(lambda: dp.run_prepacked(C), C, args.steps)
        if kern is None:
            kern = kernel_only(dp, C, args.steps)
    if kern is None:
        kern = call
    cold, cold_med = cold_times(dp, A, B, C)
    flops = c.flops()
    bytes_ = A.element_size() * (A.numel() + B.numel() + C.numel())
    res = {
        "workload": str(c), "prec": prec, "M": c.problem.M, "N": c.problem.N, "K": c.problem.K,
        "plan": best.describe(), "tuned_candidates": len(ms), "tune_seconds": round(tune_s, 2),
        "wall_ms_per_step": wall / args.steps * 1e3, "warmup_calls": nwarm, "call_ms": call * 1e3,
        "kernel_ms": kern * 1e3,
        "cold_min_ms": cold * 1e3, "cold_median_ms": cold_med * 1e3,
        "gflops": flops / (wall / args.steps) / 1e9, "kernel_gflops": flops / kern / 1e9,
        "kernel_gbs": bytes_ / kern / 1e9, "algorithmic_flops": flops, "algorithmic_bytes": bytes_,
    }
    if prec != "f64" and wl in ("abcd-aebf-dfce", "abcdef-dgfa-gceb", "abc-akb-kc"):
        C.zero_()
        dp.run(A, B, C)
        torch.cuda.synchronize()
        ref = fp64_reference(c, A, B)
        res["rel_err_vs_fp64"] = rel_err(C, ref)
        del ref
    elif prec == "f64":
        ref = fp64_reference(c, A, B)
        res["rel_err_vs_fp64_einsum"] = rel_err(C, ref)
        del ref
    log("%-7s %-18s %9.1f GFLOP/s (kernel %.3f ms, call %.3f ms, cold %.3f ms) %s" % (
        prec, wl, res["kernel_gflops"], res["kernel_ms"], res["call_ms"], res["cold_min_ms"], best.describe()))
    del A, B, C
    torch.cuda.empty_cache()
    return res


def run_batch(args, log):
    """512 FP32 contractions from the 4 batch patterns, every extent uniform over
    multiples of 8 in [8, 96] (seed 0), all in one grouped packed launch (a
    pack launch and a GETT launch over every problem), its tile and work
    order chosen by timing the model's best-ranked choices; one
    HIP-graph-captured step."""
    import numpy as np
    import torch
    from tccg_amd.batch import GettBatch, autotune_group, random_batch
    from tccg_amd.plan import F32, PACKED_TILES
    cs = random_batch(512, seed=0)
    rng = np.random.default_rng(0)
    As, Bs, Cs = [], [], []
    for c in cs:
        sa, sb, sc = c.shapes()
        As.append(torch.from_numpy(rng.uniform(-1, 1, sa).astype(np.float32)).cuda())
        Bs.append(torch.from_numpy(rng.uniform(-1, 1, sb).astype(np.float32)).cuda())
        Cs.append(torch.zeros(sc, device="cuda"))
    t0 = time.perf_counter()
    tile, order, _ = autotune_group(cs, As, Bs, Cs, top=6, reps=2, verbose=args.verbose)
    tune_s = time.perf_counter() - t0
    b = GettBatch(cs, group_tile=tile, group_order=order).bind(As, Bs, Cs)
    graph = b.capture()
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(s)
        graph.replay()
        e1.record(s)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    kern = sum(e0.elapsed_time(e1) for e0, e1 in evs) / args.steps * 1e-3
    errs = []
    for i in range(0, len(cs), 32):   # every 32nd problem against FP64 einsum on the GPU
        ref = fp64_reference(cs[i], As[i], Bs[i])
        errs.append(rel_err(Cs[i], ref))
    flops = b.flops()
    grp = b._bound["group"]._bound
    res = {"workload": "512 FP32 contractions, 4 patterns, extents U{8..96 step 8}, seed 0",
           "problems": len(cs), "launches_per_step": b.launches(), "tile": list(PACKED_TILES[F32][tile]),
           "work_order": order, "work_items": grp["nwork"], "pack_jobs": grp["njobs"],
           "tune_seconds": round(tune_s, 2),
           "wall_ms_per_step": wall * 1e3, "kernel_ms": kern * 1e3, "gflops": flops / wall / 1e9,
           "kernel_gflops": flops / kern / 1e9, "algorithmic_flops": flops,
           "max_rel_err_vs_fp64_sampled": max(errs)}
    log("batch   512 problems        %9.1f GFLOP/s (graph replay %.3f ms, %d launches, tile %s, order %s)" % (
        res["kernel_gflops"], res["kernel_ms"], res["launches_per_step"], res["tile"], order))
    del As, Bs, Cs, b, graph
    torch.cuda.empty_cache()
    return res


def run_ttgt(args, log):
    """TTGT (the reference's transpose-transpose-GEMM-transpose method) on the
    compute-bound contraction in FP32: the model's best candidate as a whole
    call, and each of its transpositions alone against the HBM peak."""
    import torch
    from tccg_amd import Contraction
    from tccg_amd.autotune import time_plan
    from tccg_amd.plan import F32
    from tccg_amd.ttgt import DeviceTtgt, best_candidates, ttgt_candidates
    from tccg_amd.workloads import get
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    cand = best_candidates(ttgt_candidates(c.problem, 4), 1)[0]
    d = DeviceTtgt(c, cand, F32, tune=True)
    sa, sb, sc = c.shapes()
    A, B = seeded_inputs(c, torch.float32)
    C = torch.zeros(sc, dtype=torch.float32, device="cuda")
    t = time_plan(d, A, B, C, 1.0, 0.0, reps=max(3, args.steps // 4))
    C.zero_()
    d.run(A, B, C)
    torch.cuda.synchronize()
    err = rel_err(C, fp64_reference(c, A, B))
    moves = []
    s = torch.cuda.current_stream()
    for name, tr, src in (("A", d.tA, A), ("B", d.tB, B)):
        if tr is None:
            continue
        n = tr.plan.elements()
        dst = torch.empty(n, dtype=torch.float32, device="cuda")
        tr.run(src.reshape(-1), dst)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(args.steps):
            tr.run(src.reshape(-1), dst)
        e1.record(s)
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) / args.steps * 1e3
        gbs = tr.plan.moved_bytes(4) / us / 1e3
        moves.append({"operand": name, "plan": tr.plan.describe(), "us": round(us, 2), "gbs": round(gbs, 1),
                      "hbm_frac": round(gbs / PEAK_HBM_GBS, 4)})
        log("ttgt    transpose %s %8.1f us %7.0f GB/s (%.0f%% of 8 TB/s) %s"
            % (name, us, gbs, 100 * gbs / PEAK_HBM_GBS, tr.plan.describe()))
    res = {"workload": str(c), "candidate": "ttgemmt_" + cand.name, "gemm_plan": d.gemm_plan.describe(),
           "call_ms": t * 1e3, "gflops": c.flops() / t / 1e9, "rel_err_vs_fp64": err, "transposes": moves,
           "workspace_bytes": d.workspace_bytes()}
    log("ttgt    %-18s %9.1f GFLOP/s (call %.3f ms) %s, GEMM %s" % ("abcd-aebf-dfce", res["gflops"], t * 1e3,
                                                                res["candidate"], res["gemm_plan"]))
    return res


def run_mfma(args, log):
    """Sustained MFMA rates of this chip (tccg_amd/csrc/mfma_probe.hip: every
    SIMD busy with independent register-only MFMAs): FP32 32x32x2, BF16
    32x32x16 and FP64 16x16x4.  A 32x32x16 block of a contraction takes 8 FP32
    MFMAs or 6 BF16 ones (the three-way split), so (BF16 rate / 6) / FP32 rate
    bounds the split-BF16 / FP32 throughput ratio of any GETT kernel."""
    import ctypes
    import torch
    from tccg_amd import build as hipbuild
    path = hipbuild.probe_library_path()
    if not path.exists():
        log("mfma probe library missing (build() compiles it)")
        return None
    lib = ctypes.CDLL(str(path))
    lib.tccg_probe_mfma.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.tccg_probe_mfma_flops.restype = ctypes.c_double
    lib.tccg_probe_mfma_round.argtypes = [ctypes.c_int]
    s = torch.cuda.current_stream()
    res = {}
    # (kind, name, iterations, MFMAs per wave and iteration)
    for kind, name, iters, per_iter in ((0, "f32_32x32x2", 1000, 8), (1, "bf16_32x32x16", 2000, 8),
                                        (2, "f64_16x16x4", 1000, 8), (3, "bf16x3_gett_pattern", 300, 54),
                                        (4, "f32_gett_pattern", 120, 72)):
        rnd = lib.tccg_probe_mfma_round(kind)
        if rnd <= 0:
            raise RuntimeError("mfma probe: no resident workgroups")
        blocks = 4 * rnd   # four whole rounds: every SIMD equally busy to the end
        out = torch.empty(blocks * 256, dtype=torch.float32, device="cuda")

        def launch():
            if lib.tccg_probe_mfma(kind, blocks, iters, ctypes.c_void_p(out.data_ptr()),
                                   ctypes.c_void_p(s.cuda_stream)) != 0:
                raise RuntimeError("mfma probe launch failed")
        launch()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(5):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            launch()
            e1.record(s)
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        flops = lib.tccg_probe_mfma_flops(kind) * blocks * 4 * iters * per_iter
        res[name + "_tflops"] = round(flops / best / 1e12, 1)
    res["split_bf16_ceiling_ratio"] = round(res["bf16_32x32x16_tflops"] / 6 / res["f32_32x32x2_tflops"], 3)
    res["split_bf16_ceiling_ratio_gett_pattern"] = round(
        res["bf16x3_gett_pattern_tflops"] / 6 / res["f32_gett_pattern_tflops"], 3)
    log("mfma    sustained TFLOP/s: FP32 %.1f, BF16 %.1f, FP64 %.1f; split-BF16 / FP32 ceiling %.2fx" % (
        res["f32_32x32x2_tflops"], res["bf16_32x32x16_tflops"], res["f64_16x16x4_tflops"],
        res["split_bf16_ceiling_ratio"]))
    log("mfma    in the GETT's operand pattern (3 x 3 fragments, 96 x 96 per wave): FP32 %.1f, split-BF16 "
        "products %.1f TFLOP/s; ceiling %.2fx" % (res["f32_gett_pattern_tflops"], res["bf16x3_gett_pattern_tflops"],
                                                 res["split_bf16_ceiling_ratio_gett_pattern"]))
    return res


def run_log(args, log):
    """LoG (the reference's loop-over-GEMM method: every instance of the
    single-index GEMM one problem of a grouped packed launch, the contracted
    loops folded into k) on the compute-bound contraction, the model's best
    candidate, in FP32 and FP64."""
    import torch
    from tccg_amd import Contraction
    from tccg_amd.autotune import time_plan
    from tccg_amd.loop_gemm import DeviceLoG, best_log_candidates, log_candidates
    from tccg_amd.plan import F32, F64, PACKED_TILES, PREC_NAMES
    from tccg_amd.workloads import get
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    cand = best_log_candidates(log_candidates(c.problem), 1)[0]
    out = {}
    for prec, dt in ((F32, torch.float32), (F64, torch.float64)):
        A, B = seeded_inputs(c, dt)
        C = torch.zeros(c.C0.torch_shape(), dtype=dt, device="cuda")
        d = DeviceLoG(c, cand, A, B, C, 1.0, 0.0, prec=prec)
        t = time_plan(d, A, B, C, reps=3)
        C.zero_()
        d.launch()
        torch.cuda.synchronize()
        err = rel_err(C, fp64_reference(c, A, B))
        lay = d.layout
        name = PREC_NAMES[prec]
        out[name] = {"workload": str(c), "candidate": "loopGemm_" + cand.name, "launches": d.nlaunches,
                     "gemm_instances": len(lay.free_bases), "k_folded": lay.K,
                     "tile": list(PACKED_TILES[prec][lay.tile]), "call_ms": t * 1e3, "gflops": c.flops() / t / 1e9,
                     "rel_err_vs_fp64": err}
        log("log     %-7s %-18s %9.1f GFLOP/s (call %.3f ms, %d launches, %d GEMMs of k %d) %s" % (
            name, "abcd-aebf-dfce", out[name]["gflops"], t * 1e3, d.nlaunches, len(lay.free_bases), lay.K,
            out[name]["candidate"]))
        del A, B, C, d
    return out


def run_grad(args, log):
    """Forward + backward (two adjoint GETTs) of C[a,b,c,d] = A[a,k,l,b] * B[l,c,k,d],
    72/32, loss 0.5*||C - T||^2, FP32; gradients against FP64 torch.einsum autograd."""
    import torch
    from tccg_amd import Contraction
    from tccg_amd.autograd import adjoints, contract
    from tccg_amd.workloads import get
    c = Contraction.from_tccg(get("abcd-aklb-lckd"))
    g = torch.Generator(device="cuda").manual_seed(0)
    A = torch.randn(c.A0.torch_shape(), device="cuda", generator=g)
    B = torch.randn(c.B0.torch_shape(), device="cuda", generator=g)
    T = torch.randn(c.C0.torch_shape(), device="cuda", generator=g)
    out = {}
    for prec in ("f32", "bf16x3", "bf16"):
        A1, B1 = A.clone().requires_grad_(), B.clone().requires_grad_()

        def step():
            A1.grad = None
            B1.grad = None
            L = 0.5 * ((contract(A1, B1, c, prec, tune=True) - T) ** 2).sum()
            L.backward()
        for _ in range(max(1, args.warmup)):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        A2, B2 = A.double().requires_grad_(), B.double().requires_grad_()
        eq = fp64_reference.__globals__["_einsum_eq"](c)
        L2 = 0.5 * ((torch.einsum(eq, A2, B2) - T.double()) ** 2).sum()
        L2.backward()
        out[prec] = {"step_ms": wall * 1e3, "gett_gflops_per_step": 3 * c.flops() / wall / 1e9,
                     "rel_err_dA": rel_err(A1.grad, A2.grad), "rel_err_dB": rel_err(B1.grad, B2.grad)}
        log("grad    %-7s abcd-aklb-lckd %8.1f GFLOP/s (3 GETTs + elementwise, %.3f ms/step) dA err %.2e" % (
            prec, out[prec]["gett_gflops_per_step"], wall * 1e3, out[prec]["rel_err_dA"]))
    return out


def profiled_traffic(workload: str, prec: str, plan: str):
    """HBM bytes per launch of the dominant kernel from profiles/traffic.json
    (calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE of the same workload,
    precision and plan); None when no profile of this exact plan is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entries = json.load(f)
    except (OSError, ValueError):
        return None
    for e in entries:
        if e.get("workload") == workload and e.get("prec") == prec and e.get("plan") == plan:
            return e.get("hbm_bytes_per_launch")
    return None


CPU_CONFIGS = ("cpu48", "abcd-aebf-dfce", "abc-akb-kc")


def cpu_baseline(args, log):
    """The oracle's AVX2 GETT restatement on the host cores: the CPU
    configuration (C[a,b,c,d] = A[a,k,c] * B[k,d,b], 48/64) and both plain FP32
    GPU configurations, each a bounded sample (the restated planner's best
    three candidates timed on all physical cores, the fastest also on one core,
    about --cpu-sample-seconds / 3 per configuration), checked against the
    FP64 einsum oracle.  `value` is the compute-bound configuration's
    all-core rate (the bench line's workload)."""
    try:
        from oracle import avx2_gett
    except Exception as e:  # the restatement is optional for the GPU line
        log("cpu baseline unavailable: %s" % e)
        return None
    per = {wl: avx2_gett.bench_workload(wl, seconds=args.cpu_sample_seconds / len(CPU_CONFIGS), log=log)
           for wl in CPU_CONFIGS}
    head = per["abcd-aebf-dfce"]
    return {"value": head["gflops"], "unit": "GFLOP/s", "cores": head["cores"], "kind": "port",
            "sample": ("whole contractions, min over repetitions after the restated planner's top-3 candidates "
                       "were timed; configurations: %s" % ", ".join(
                           "%s %.1f GFLOP/s (%d cores) / %.1f (1 core) [%s]" % (
                               wl, r["gflops"], r["cores"], r["gflops_one_core"], r["candidate"].split("_mr")[0])
                           for wl, r in per.items())),
            "one_core_value": head["gflops_one_core"], "cpu": avx2_gett.cpu_model(), "configs": per}


def main():
    args = parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the GETT backend runs one contraction per GPU and has "
              "no multi-GPU path (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    log = (lambda s: print(s, file=sys.stderr, flush=True))
    if args.copy_bytes:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copy: %d bytes read, %d bytes written" % (args.copy_bytes, args.copy_bytes))
        del src, dst
    forms = {
        "f32": ("abcd-aebf-dfce", "f32"),
        "bw": ("abc-akb-kc", "f32"),
        "f64": ("abcd-aebf-dfce", "f64"),
        "bf16x3": ("abcd-aebf-dfce", "bf16x3"),
        "bf16": ("abcd-aebf-dfce", "bf16"),
        "fp8_f32": ("abcdef-dgfa-gceb", "f32"),
        "fp8_bf16x3": ("abcdef-dgfa-gceb", "bf16x3"),
        "fp8": ("abcdef-dgfa-gceb", "fp8"),
        "batch": None,
        "grad": None,
        "ttgt": None,
        "log": None,
        "mfma": None,
    }
    if args.only:
        forms = {args.only: forms[args.only]}
    results = {}
    for key, spec in forms.items():
        if key == "batch":
            results[key] = run_batch(args, log)
        elif key == "grad":
            results[key] = run_grad(args, log)
        elif key == "ttgt":
            results[key] = run_ttgt(args, log)
        elif key == "log":
            results[key] = run_log(args, log)
        elif key == "mfma":
            results[key] = run_mfma(args, log)
        else:
            results[key] = run_form(key, spec[0], spec[1], args, log)
    head = results.get("f32") or next(iter(results.values()))
    details = {}
    if "bf16x3" in results and "f32" in results:
        details["split_bf16_speedup_vs_f32"] = results["bf16x3"]["kernel_gflops"] / results["f32"]["kernel_gflops"]
        details["split_bf16_err_ratio_vs_f32"] = (results["bf16x3"]["rel_err_vs_fp64"]
                                                   / max(results["f32"]["rel_err_vs_fp64"], 1e-30))
    if "bw" in results:
        details["bw_hbm_roofline_frac"] = results["bw"]["kernel_gbs"] / PEAK_HBM_GBS
    if results.get("mfma") and "split_bf16_speedup_vs_f32" in details:
        details["split_bf16_ceiling_ratio_measured"] = results["mfma"]["split_bf16_ceiling_ratio"]
        details["split_bf16_speedup_vs_ceiling"] = (details["split_bf16_speedup_vs_f32"]
                                                    / results["mfma"]["split_bf16_ceiling_ratio"])
    for k in ("fp8_f32", "fp8_bf16x3", "fp8"):
        if k in results:
            details["%s_rel_err_vs_fp64" % k] = results[k].get("rel_err_vs_fp64")
    cpu = None if (args.no_cpu_baseline or args.only) else cpu_baseline(args, log)
    if not head or "kernel_gflops" not in head or "prec" not in head:
        # a composite form alone (--only batch / grad / ttgt / log / mfma): its own record only
        print(json.dumps({"metric": METRIC, "form": args.only, "forms": results}), flush=True)
        return
    traffic = profiled_traffic(head["workload"], head["prec"], head["plan"])
    if head is results.get("bw"):
        # the bandwidth-bound contraction alone (--only bw): bytes against HBM
        roofline = {"bound": "hbm", "achieved": round(head["kernel_gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(head["kernel_gbs"] / PEAK_HBM_GBS, 4), "traffic": traffic}
    else:
        achieved_tf = head["kernel_gflops"] / 1e3
        roofline = {"bound": "mfma", "achieved": round(achieved_tf, 2), "peak": PEAK_TFLOPS[head["prec"]],
                    "unit": "TFLOP/s", "frac": round(achieved_tf / PEAK_TFLOPS[head["prec"]], 4), "traffic": traffic}
    # the time behind `achieved`: HIP events around graph replays of the GETT kernel alone
    roofline["time_source"] = ("hip events, graph replay of %d kernel launches: %.4f ms per launch"
                               % (args.steps, head["kernel_ms"]))
    line = {
        "metric": METRIC,
        "value": round(head["gflops"], 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        # the warmup calls made: at least `warmup`, more until the GPU had run
        # them for SETTLE_SECONDS (clocks out of their idle state)
        "warmup_calls": head["warmup_calls"],
        "ms_per_step": round(head["wall_ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head["prec"],
        "data": "synthetic: iid uniform[-1,1] from numpy default_rng(0), A then B, operands resident in HBM",
        "config": {"workload": head["workload"], "M": head["M"], "N": head["N"], "K": head["K"],
                   "alpha": 1.0, "beta": 0.0, "plan": head["plan"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "forms": results,
        "details": details,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
