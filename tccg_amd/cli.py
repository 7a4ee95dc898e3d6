"""The tccg command line with the gfx950 target.

    python -m tccg_amd [--arch=gfx950] [--floatType=s|d] [--maxImplementations=N]
                       [--testing] [--generateOnly] ... file.tccg

Flags and their meaning follow the reference's argparse table
(tccg/tccg.py:1253-1285) and TccgArgs defaults (tccg/tccg.py:30-59).  The
flow follows Tccg.codeGen (tccg/tccg.py:188-430): look the contraction up in
the tuning database, otherwise rank GETT candidates with the performance
model, time the best --maxImplementations of them on the GPU (printing
"<name>: X GFLOPS // estimated: Y GFLOPS" per candidate, tccg/tccg.py:1052-1056),
store them, and emit the fastest as tccg_implementations/gett.hip + gett.hpp
+ libgett.so exporting gett(A, B, C, alpha, beta) (tccg/tccg.py:158-186).
Unless --noLoG / --noTTGT are given, the best LoG candidates
(tccg/gemm_loop.py; tccg_amd.loop_gemm) and the TTGT candidates of minimal
transposition cost (tccg/ttgemmt.py:133-336; tccg_amd.ttgt) are timed too,
at most --maxImplementationsLoG / --maxImplementationsTTGT of each (default 1,
as in the reference), and reported as the reference's harness reports them
("loopGemm_<name>: X GFLOPS", "ttgemmt_<name>: X GFLOPS", tccg/tccg.py:1060-1174;
"Best loop-based / TTGT-based implementation ...", tccg/tccg.py:377, 399).
When LoG or TTGT beats the fastest GETT, its code is emitted instead, as the
reference does (loopOverGemm.hip / ttgemmt.hip with the reference's method
entry point beside the same gett() boundary, tccg_amd.codegen_methods).  The summary line "Best Loop/TTGT/GETT/reference/GEMM/TCL:
..." is printed as the reference does (tccg/tccg.py:426); the GEMM and TCL
methods are not part of this backend and report 0.00 / -1.

Only --arch=gfx950 is provided; the reference's CPU/CUDA architectures are
rejected.  --floatType=s runs FP32 (or, with --precision, the split-BF16 or
FP8 forms), --floatType=d runs FP64.
"""
from __future__ import annotations

import argparse
import os
import socket
import sys
import time
from pathlib import Path

from .frontend import TccgSyntaxError, contraction_str, fuse_indices, read_tccg_file
from .plan import (BF16X3, F32, F64, FP8, PREC_NAMES, UnsupportedContraction, candidate_plans, build_plan,
                   estimate_seconds, plan_from_kind, plan_kind, tile_shape_of, tile_tables_digest)

OKGREEN = "\033[92m"
FAIL = "\033[91m"
WARNING = "\033[93m"
ENDC = "\033[0m"


def make_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="Generate high-performance GPU code for a given tensor contraction.")
    p.add_argument("filename", metavar="filename", type=str,
                   help="filename of the file which contains the symbolic description of the tensor contraction")
    p.add_argument("--verbose", action="store_true", help="Verbose output (useful for debugging).")
    p.add_argument("--testing", action="store_true", help="enable testing of the generated routines.")
    p.add_argument("--keep", action="store_true", help="keep intermediate files")
    p.add_argument("--noBatchedGEMM", action="store_true",
                   help="accepted for compatibility (LoG always runs its GEMMs as one batched launch)")
    p.add_argument("--generateOnly", action="store_true", help="only generate code (no timing on the GPU)")
    p.add_argument("--gettOnly", action="store_true", help="Only use GETT implementation (always the case here)")
    p.add_argument("--noGETT", action="store_true", help="Do not use GETT implementation")
    p.add_argument("--noLoG", action="store_true", help="do not time the LoG method")
    p.add_argument("--noGEMM", action="store_true", help="accepted for compatibility")
    p.add_argument("--noTTGT", action="store_true", help="do not time the TTGT method")
    p.add_argument("--useTCL", action="store_true", help="not supported by this backend")
    p.add_argument("--useDynamicMemory", action="store_true", help="accepted for compatibility")
    p.add_argument("--timePacking", action="store_true", help="accepted for compatibility")
    p.add_argument("--ignoreDatabase", action="store_true",
                   help="ignore the tuning database (no lookup of an existing solution)")
    p.add_argument("--fastMeasurements", action="store_true", help="time fewer repetitions per candidate")
    p.add_argument("--numThreads", type=int, help="accepted for compatibility (one GPU per contraction)")
    p.add_argument("--maxImplementations", type=int,
                   help="maximum number of GETT candidates to evaluate (default: 16). -1 denotes infinity.")
    p.add_argument("--maxImplementationsTTGT", type=int,
                   help="maximum number of TTGT candidates to evaluate (default: 1). -1 denotes infinity.")
    p.add_argument("--maxImplementationsLoG", type=int,
                   help="maximum number of LoG candidates to evaluate (default: 1). -1 denotes infinity.")
    p.add_argument("--maxWorkspace", type=float, help="accepted for compatibility")
    p.add_argument("--arch", metavar="arch", type=str, help="architecture: gfx950 (default)")
    p.add_argument("--compiler", metavar="compiler", type=str, help="accepted for compatibility (hipcc is used)")
    p.add_argument("--floatType", metavar="floatType", type=str, help="floatType can be either 's' or 'd'.")
    p.add_argument("--precision", choices=["fp32", "split-bf16", "fp8"], default="fp32",
                   help="arithmetic of --floatType=s: plain FP32 MFMA, three-way BF16 split, or FP8 e4m3")
    p.add_argument("--affinity", metavar="affinity", type=str, help="accepted for compatibility")
    p.add_argument("--output", type=str, default=None,
                   help="output directory (default: ./tccg_implementations)")
    p.add_argument("--database", type=str, default=None, help="tuning database path (default: $TCCG_DB or "
                                                               "~/.cache/tccg_amd/tccg.db)")
    return p


def settings_str(args, prec: int) -> str:
    rows = [("Version", "tccg_amd 0.1.0 (gfx950)"), ("inputfile", args.filename),
            ("maxImplementations", str(args.maxImplementations)), ("compiler", "hipcc"),
            ("architecture", args.arch), ("floatType", {F64: "double"}.get(prec, "float")),
            ("precision", PREC_NAMES[prec]), ("hostname", socket.gethostname())]
    out = "-----------------SETTINGS-----------------\n"
    out += "".join(k.ljust(20) + v + "\n" for k, v in rows)
    return out + "------------------------------------------"


def candidate_name(plan) -> str:
    """Reference-style candidate name (tccg/gett.py:578-579); decodable by the
    reference's decodeName (tccg/gett.py:561-576)."""
    from .plan import WAVES, plan_kind
    pb = plan.problem
    blocked = 0 if plan.lin_m.is_plain() else 1
    # var: bit 0 blocked m linearisation, higher bits the plan variant
    var = blocked | (plan_kind(plan) << 1)
    WM, WN = (2, 2) if plan.packed else WAVES.get((plan.prec, plan.tile), (2, 2))
    wtm, wtn = plan.BM // WM, plan.BN // WN
    sub = 16 if plan.prec in (F64, FP8) else 32
    return "gett_var%d_mc%d_nc%d_kc%d_mc1%d_nc1%d_mr%d_nr%d_%s_%s_%s" % (
        var, plan.BM, plan.BN, plan.BK, wtm, wtn, sub, sub, pb.opA.index_str(), pb.opB.index_str(),
        pb.C.index_str())


def _precision(args) -> int:
    ft = args.floatType or "s"
    if ft == "d":
        return F64
    if ft != "s":
        raise SystemExit(FAIL + "ERROR: floatType %s is not supported yet." % ft + ENDC)
    return {"fp32": F32, "split-bf16": BF16X3, "fp8": FP8}[args.precision]


def main(argv=None) -> int:
    args = make_parser().parse_args(argv)
    args.arch = args.arch or "gfx950"
    # defaults of TccgArgs (tccg/tccg.py:40-42); -1 denotes infinity
    for name, default in (("maxImplementations", 16), ("maxImplementationsTTGT", 1), ("maxImplementationsLoG", 1)):
        v = getattr(args, name)
        setattr(args, name, default if v is None else (1000000000 if v == -1 else v))
    if args.arch != "gfx950":
        print(FAIL + "[GETT] Error: architecture %s is not provided by this backend (use --arch=gfx950)"
              % args.arch + ENDC, file=sys.stderr)
        return 1
    if args.noGETT:
        print(FAIL + "ERROR: --noGETT leaves nothing to emit: the emitted entry point of this backend is a GETT"
              + ENDC, file=sys.stderr)
        return 1
    try:
        prec = _precision(args)
    except SystemExit as e:
        print(e, file=sys.stderr)
        return 1
    print(settings_str(args, prec))
    path = args.filename if os.path.isabs(args.filename) else os.path.join(os.getcwd(), args.filename)
    try:
        text = Path(path).read_text() if os.path.isfile(path) else None
        if text is None:
            raise TccgSyntaxError("File: %s does not exist." % path)
        from .contraction import Contraction
        contraction = Contraction.from_tccg(text)
    except (TccgSyntaxError, UnsupportedContraction) as e:
        print(FAIL + str(e) + ENDC, file=sys.stderr)
        return 1
    c0 = contraction_str(contraction.C0, contraction.A0, contraction.B0, contraction.alpha, contraction.beta)
    c1 = str(contraction)
    if c0 != c1:
        print("Contration before fusing indices:")
        print(c0)
        print("Contration after fusing indices:")
    print(c1)
    pb = contraction.problem
    print("m: %d\nn: %d\nk: %d" % (pb.M, pb.N, pb.K))

    out_dir = Path(args.output) if args.output else Path(os.getcwd()) / "tccg_implementations"
    plans = candidate_plans(pb, prec)[: max(1, args.maxImplementations)]
    print("Total amount of GETT implementations: %d" % len(plans))
    chosen = plans[0]
    best_gflops = -1.0
    ref_gflops = -1.0
    ttgt_gflops, ttgt_cand, ttgt_plan = 0.0, None, None
    log_gflops, log_cand = 0.0, None
    gpu = False
    if not args.generateOnly:
        try:
            import torch
            gpu = torch.cuda.is_available()
        except Exception:
            gpu = False
        if not gpu:
            print(WARNING + "WARNING: no GPU visible; emitting the model's best candidate without timing" + ENDC)
    if gpu:
        from .tuning_db import TuningDB
        db = TuningDB(args.database)
        device = "gfx950/" + tile_tables_digest()
        hit = None if args.ignoreDatabase else db.fastest(contraction, device, PREC_NAMES[prec], contraction.beta)
        if hit is not None and tile_shape_of(prec, hit[1], hit[3]) != hit[5]:
            print(WARNING + "[database] entry %s names a tile of another shape; re-tuning" % hit[0] + ENDC)
            hit = None
        if hit is not None:
            name, tile, blocked, kind, gf, _shape = hit
            chosen = plan_from_kind(pb, prec, tile, blocked, kind)
            best_gflops = gf
            print(OKGREEN + "[database] reusing %s (%.2f GFLOPS)" % (name, gf) + ENDC)
        else:
            best_gflops, chosen, ref_gflops = _time_candidates(args, contraction, prec, plans, db, device)
            if chosen is None:
                return 1
            if not args.noLoG:
                log_gflops, log_cand = _time_log(args, contraction, prec)
                if log_gflops is None:
                    return 1
            if not args.noTTGT:
                ttgt_gflops, ttgt_cand, ttgt_plan = _time_ttgt(args, contraction, prec)
                if ttgt_gflops is None:
                    return 1
        db.close()
    # emitFinalCode (tccg/tccg.py:158-186): the fastest method's code; GETT
    # wins ties, then LoG over TTGT
    method = "gett"
    if prec in (F32, F64) and best_gflops > 0:
        if log_cand is not None and log_gflops > best_gflops and log_gflops >= ttgt_gflops:
            method = "log"
        elif ttgt_cand is not None and ttgt_gflops > best_gflops:
            method = "ttgt"
    try:
        if method == "gett":
            print(OKGREEN + "[generate] emitting %s" % candidate_name(chosen) + ENDC)
            from . import codegen
            res = codegen.generate(chosen, text, out_dir, "gett")
        else:
            from . import codegen_methods
            cand = log_cand if method == "log" else ttgt_cand
            label = ("loopGemm_" if method == "log" else "ttgemmt_") + cand.name
            print(OKGREEN + "[generate] emitting %s (faster than the best GETT)" % label + ENDC)
            res = codegen_methods.generate_method(method, contraction, cand, prec, text, out_dir,
                                                  gemm_plan=ttgt_plan if method == "ttgt" else None)
    except Exception as e:
        print(FAIL + "[TCC Error] compilation failed: %s" % e + ENDC, file=sys.stderr)
        return 1
    print("The generated code is available at:", res["source"])
    print("Best gett-based implementation (%s) attained: %.2f GFLOPS/s" % (candidate_name(chosen), max(best_gflops, 0)))
    print("Best Loop/TTGT/GETT/reference/GEMM/TCL: %.2f / %.2f / %.2f / %.2f / %.2f / %.2f GFLOPS"
          % (max(log_gflops, 0.0), max(ttgt_gflops, 0.0), max(best_gflops, 0.0), ref_gflops, 0.0, -1.0))
    return 0


def harness_inputs(n: int, which: str, dtype, device="cuda"):
    """The generated harness's deterministic tensor contents
    (tccg/tccg.py:957-971): A[i] = (10 + (i+1)*17 % 100) / 110,
    B[i] = (10 + (i+2)*3 % 100) / 110, C[i] = (10 + (i+3)*13 % 100) / 110,
    the integer part computed exactly, rounded once to FP32 ((float) cast)
    and divided in double, then stored in the tensor's type."""
    import torch
    mul, add = {"A": (17, 1), "B": (3, 2), "C": (13, 3)}[which]
    i = torch.arange(n, dtype=torch.int64, device=device)
    v = (10 + ((i + add) * mul) % 100).to(torch.float32).to(torch.float64) / 110.0
    return v.to(dtype)


def harness_equal(x, ref, is_double: bool) -> bool:
    """The generated harness's check (tccg/tccg.py:843-881): for every element
    min(|d| / max(|a|, |b|), |d|) <= 1e-4 (float) or 1e-9 (double), where
    d = |a| - |b| (the harness compares absolute values)."""
    import torch
    a = x.double().abs().reshape(-1)
    b = ref.double().abs().reshape(-1)
    d = (a - b).abs()
    rel = torch.where(d > 0, torch.minimum(d / torch.maximum(a, b), d), torch.zeros_like(d))
    return bool((rel <= (1e-9 if is_double else 1e-4)).all())


def _test_inputs(contraction, prec, with_ref: bool):
    """A, B, C in the generated harness's deterministic contents, and the FP64
    result alpha * A.B + beta * C when with_ref."""
    import torch
    dt = torch.float64 if prec == F64 else torch.float32
    sa, sb, sc = contraction.shapes()
    n = lambda s: int(__import__("numpy").prod(s))
    A = harness_inputs(n(sa), "A", dt).reshape(sa)
    B = harness_inputs(n(sb), "B", dt).reshape(sb)
    C0 = harness_inputs(n(sc), "C", dt).reshape(sc)
    ref = _fp64_reference(contraction, A, B, C0) if with_ref else None
    return A, B, C0, ref


def _check_result(name: str, prec: int, C, ref, tol: float) -> bool:
    """--testing: the reference harness's element check for FP32 / FP64
    results and the relative Frobenius bar of this backend (both must pass;
    the reduced-precision forms only the latter)."""
    err = float((C.double() - ref).norm() / ref.norm())
    ok = err <= tol
    if prec in (F32, F64):
        ok = ok and harness_equal(C, ref, prec == F64)
    if not ok:
        print(FAIL + "ERROR: version '%s' failed (relative error %.3e)" % (name, err) + ENDC)
    return ok


def _time_log(args, contraction, prec):
    """Time the LoG candidates with the largest GEMM and batch (at most
    --maxImplementationsLoG; the grouped launch runs FP32, FP64, split-BF16
    and BF16).  Best GFLOP/s (0 when LoG does not apply), or None if
    --testing found a wrong result."""
    import torch
    from .autotune import time_plan
    from .group import GROUP_PRECS
    from .loop_gemm import DeviceLoG, best_log_candidates, log_candidates
    cands = log_candidates(contraction.problem)
    if cands is None or prec not in GROUP_PRECS:
        print("LoG not possible" if cands is None else "LoG: not run in %s" % PREC_NAMES[prec])
        return 0.0, None
    best = best_log_candidates(cands, args.maxImplementationsLoG)
    print("Total amount of LoG implementations: %d" % len(cands))
    print("Best LoG candidates: %s" % ", ".join(c.name for c in best))
    A, B, C0, ref = _test_inputs(contraction, prec, args.testing)
    C = C0.clone()
    top, top_name, top_cand = -1.0, "", None
    for cand in best:
        d = DeviceLoG(contraction, cand, A, B, C, prec=prec)
        name = "loopGemm_" + cand.name
        if ref is not None:
            C.copy_(C0)
            d.launch()
            torch.cuda.synchronize()
            if not _check_result(name, prec, C, ref, 1e-5):
                print(FAIL + "[TCC Error] runtime error." + ENDC, file=sys.stderr)
                return None, None
        C.copy_(C0)
        t = time_plan(d, A, B, C, contraction.alpha, contraction.beta, reps=2 if args.fastMeasurements else 5)
        gf = contraction.flops() / t / 1e9
        print("%s: %.2f GFLOPS" % (name, gf))
        if gf > top:
            top, top_name, top_cand = gf, name, cand
    print("Best loop-based implementation (%s) attained: %.2f GFLOPS/s" % (top_name, max(top, 0.0)))
    return top, top_cand


def _time_ttgt(args, contraction, prec):
    """Time the TTGT candidates of minimal model cost (at most
    --maxImplementations); best GFLOP/s, or None if --testing found a wrong
    result."""
    import torch
    from .autotune import time_plan
    from .ttgt import DeviceTtgt, best_candidates, ttgt_candidates
    esz = 8 if prec == F64 else 4
    cands = ttgt_candidates(contraction.problem, esz)
    best = best_candidates(cands, args.maxImplementationsTTGT)
    print("Total amount of TTGT implementations: %d" % len(cands))
    print("Best TTGT candidates: %s" % ", ".join(c.name for c in best))
    A, B, C0, ref = _test_inputs(contraction, prec, args.testing)
    C = C0.clone()
    tol = {F32: 1e-5, F64: 1e-13, BF16X3: 1e-5, FP8: 8e-2}.get(prec, 1e-2)
    top, top_name, top_cand, top_plan = -1.0, "", None, None
    for cand in best:
        dt_ = DeviceTtgt(contraction, cand, prec, tune=True)
        name = "ttgemmt_" + cand.name
        if ref is not None:
            C.copy_(C0)
            dt_.run(A, B, C)
            torch.cuda.synchronize()
            if not _check_result(name, prec, C, ref, tol):
                print(FAIL + "[TCC Error] runtime error." + ENDC, file=sys.stderr)
                return None, None, None
        C.copy_(C0)
        t = time_plan(dt_, A, B, C, contraction.alpha, contraction.beta, reps=2 if args.fastMeasurements else 5)
        gf = contraction.flops() / t / 1e9
        print("%s: %.2f GFLOPS" % (name, gf))
        if gf > top:
            top, top_name, top_cand, top_plan = gf, name, cand, dt_.gemm_plan
        del dt_
    print("Best TTGT-based implementation (%s) attained: %.2f GFLOPS/s" % (top_name, max(top, 0.0)))
    return top, top_cand, top_plan


def _time_candidates(args, contraction, prec, plans, db, device):
    import numpy as np
    import torch
    from .autotune import time_plan
    from .runtime import DevicePlan
    A, B, C0, _ = _test_inputs(contraction, prec, False)
    C = C0.clone()
    ref = None
    ref_gflops = -1.0
    if args.testing:
        t0 = time.perf_counter()
        ref = _fp64_reference(contraction, A, B, C0)
        ref_gflops = contraction.flops() / (time.perf_counter() - t0) / 1e9
    print(OKGREEN + "[running] Measuring the runtime of all variants ..." + ENDC)
    mid = db.measurement_id(contraction, device, PREC_NAMES[prec], contraction.beta, create=True)
    best = (-1.0, None)
    reps = 2 if args.fastMeasurements else 5
    tol = {F32: 1e-5, F64: 1e-13, BF16X3: 1e-5, FP8: 8e-2}[prec]
    failed = 0
    for p in plans:
        dp = DevicePlan(p)
        name = candidate_name(p)
        if ref is not None:
            C.copy_(C0)
            dp.run(A, B, C, contraction.alpha, contraction.beta)
            torch.cuda.synchronize()
            if not _check_result(name, prec, C, ref, tol):
                failed += 1
                continue
        C.copy_(C0)
        t = time_plan(dp, A, B, C, contraction.alpha, contraction.beta, reps=reps)
        gf = contraction.flops() / t / 1e9
        est = contraction.flops() / estimate_seconds(p, contraction.beta) / 1e9
        print("%s: %.2f GFLOPS // estimated: %.2f GFLOPS" % (name, gf, est))
        db.insert_gett(mid, name, p.tile, not p.lin_m.is_plain(), gf, est, plan_kind(p), (p.BM, p.BN, p.BK))
        if gf > best[0]:
            best = (gf, p)
    if failed:
        print(FAIL + "[TCC Error] runtime error." + ENDC, file=sys.stderr)
        return best[0], None, ref_gflops
    return best[0], best[1], ref_gflops


def _fp64_reference(contraction, A, B, C0):
    import string
    import torch
    letters = {}
    pool = iter(string.ascii_letters)
    for lab in [i.label for i in contraction.C0.indices + contraction.A0.indices + contraction.B0.indices]:
        letters.setdefault(lab, next(pool))
    rev = lambda T: "".join(letters[i.label] for i in reversed(T.indices))
    eq = "%s,%s->%s" % (rev(contraction.A0), rev(contraction.B0), rev(contraction.C0))
    out = torch.einsum(eq, A.double().contiguous(), B.double().contiguous()) * contraction.alpha
    if contraction.beta != 0:
        out = out + contraction.beta * C0.double()
    torch.cuda.synchronize()
    return out


if __name__ == "__main__":
    sys.exit(main())
