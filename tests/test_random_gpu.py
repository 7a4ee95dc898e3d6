"""Seeded random contractions through every candidate plan, against the FP64
einsum oracle (FP8: against its e4m3 emulation).

Each case draws the index sets (free indices of A only, of B only,
contracted ones; any of them may be empty as long as A, B and C keep an
index), the order of every tensor's indices, ragged extents and alpha/beta,
then runs every plan the planner offers for it (gather, blocked, persistent
and streaming walks, vector kernel, packed tiles, persistent, store-wave,
split-k and stream-k plans) in every precision (tests/random_cases.py;
tests/test_plan_model.py checks the same cases' tables on the CPU).
"""
import numpy as np
import pytest

from tccg_amd.plan import BF16, BF16X3, F32, F64, FP8, candidate_plans

from .gpu_helpers import TOL, inputs, run_gpu
from .random_cases import random_contraction

pytestmark = pytest.mark.gpu

TOL_BF16 = 1e-2   # BF16 against the FP64 einsum: one BF16 product per term
TOL_FP8 = 1e-4    # FP8 against its emulation (the MFMA's group sums, about 1e-5)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3, BF16, FP8])
@pytest.mark.parametrize("seed", range(48))
def test_random_contraction_every_plan(seed, prec):
    c = random_contraction(seed)
    alpha, beta = [(1.0, 0.0), (2.4, 1.1), (-0.5, 0.0)][seed % 3]
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, seed, "normal", dtype, with_c=beta != 0)
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from oracle.precision_ref import fp8_contract
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    if prec == FP8:
        # the e4m3 emulation (per-tensor scales, exact products, FP64 sum)
        ref = fp8_contract(*labels, A, B, alpha) + (beta * C if beta != 0 else 0.0)
    else:
        ref = contract_ref(*labels, A, B, C, alpha, beta)
    plans = candidate_plans(c.problem, prec)
    assert plans
    worst = 0.0
    for i, plan in enumerate(plans):
        print("random", seed, prec, c, i, plan.describe(), flush=True)
        out = run_gpu(c, A, B, C, prec, alpha, beta, plan=plan)
        err = rel_frobenius(out, ref)
        limit = TOL_BF16 if prec == BF16 else TOL_FP8 if prec == FP8 else TOL[prec]
        assert np.isfinite(err) and err <= limit, "plan %d (%s): rel. error %.3e" % (i, plan.describe(), err)
        worst = max(worst, err)
    print("random", seed, prec, len(plans), "plans, worst rel. error %.2e" % worst, flush=True)


def _tensors(c, seed, dtype, with_c):
    import torch
    A, B, C = inputs(c, seed, "normal", dtype, with_c=with_c)
    dev = torch.device("cuda", 0)
    return A, B, C, torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("seed", range(48))
def test_random_contraction_ttgt(seed, prec):
    """TTGT (transposition kernels around the GETT) on the random cases: the
    model's best candidates plus one that transposes C and one that exchanges
    A and B, with beta == 0 and beta != 0 (candidate bookkeeping is modelled
    on the CPU for every candidate by tests/test_ttgt.py's restatement)."""
    import torch
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.ttgt import DeviceTtgt, best_candidates, ttgt_candidates
    c = random_contraction(seed)
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C, tA, tB = _tensors(c, seed, dtype, True)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    cands = ttgt_candidates(c.problem, 8 if prec == F64 else 4)
    chosen = best_candidates(cands, 3)
    for pred in (lambda x: x.transC, lambda x: x.swap == 1):
        extra = next((x for x in cands if pred(x)), None)
        if extra is not None and extra not in chosen:
            chosen.append(extra)
    for cand in chosen:
        d = DeviceTtgt(c, cand, prec)
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            print("random ttgt", seed, prec, c, cand.name, alpha, beta, flush=True)
            tC = torch.from_numpy(C).to(tA.device)
            d.run(tA, tB, tC, alpha, beta)
            torch.cuda.synchronize()
            ref = contract_ref(*labels, A, B, C, alpha, beta)
            assert rel_frobenius(tC.cpu().numpy(), ref) <= TOL[prec], (cand.name, alpha, beta)


def _log_seeds():
    from tccg_amd.loop_gemm import log_candidates
    return [s for s in range(48) if log_candidates(random_contraction(s).problem)]


@pytest.mark.parametrize("seed", _log_seeds())
def test_random_contraction_log(seed):
    """LoG (every GEMM instance one problem of a grouped launch) on the random
    cases that have LoG candidates, every candidate, FP32 and FP64."""
    import torch
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.loop_gemm import DeviceLoG, log_candidates
    c = random_contraction(seed)
    cands = log_candidates(c.problem)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for prec in (F32, F64):
        dtype = np.float64 if prec == F64 else np.float32
        A, B, C, tA, tB = _tensors(c, seed, dtype, True)
        for cand in cands:
            for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
                print("random log", seed, prec, c, cand.name, alpha, beta, flush=True)
                tC = torch.from_numpy(C).to(tA.device)
                d = DeviceLoG(c, cand, tA, tB, tC, alpha, beta, prec=prec)
                d.launch()
                torch.cuda.synchronize()
                ref = contract_ref(*labels, A, B, C, alpha, beta)
                assert rel_frobenius(tC.cpu().numpy(), ref) <= TOL[prec], (cand.name, prec, alpha, beta)


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("seed", range(48))
def test_random_contraction_autograd(seed, prec):
    """The differentiable GETT on the random cases: the forward and both
    adjoint GETTs (whose index sets may be empty where the forward's are not)
    against torch.einsum's FP64 autograd on the CPU."""
    import torch
    from oracle.einsum_ref import rel_frobenius
    from tccg_amd.autograd import contract
    from .test_autograd_gpu import einsum_eq
    c = random_contraction(seed)
    g = torch.Generator().manual_seed(seed)
    A64 = torch.randn(c.A0.torch_shape(), dtype=torch.float64, generator=g)
    B64 = torch.randn(c.B0.torch_shape(), dtype=torch.float64, generator=g)
    dC64 = torch.randn(c.C0.torch_shape(), dtype=torch.float64, generator=g)
    Ar, Br = A64.clone().requires_grad_(), B64.clone().requires_grad_()
    Cr = c.alpha * torch.einsum(einsum_eq(c), Ar, Br)
    Cr.backward(dC64)
    dt = torch.float64 if prec == "f64" else torch.float32
    A = A64.to(dt).cuda().requires_grad_()
    B = B64.to(dt).cuda().requires_grad_()
    print("random autograd", seed, prec, c, flush=True)
    C = contract(A, B, c, prec)
    C.backward(dC64.to(dt).cuda())
    torch.cuda.synchronize()
    tol = 1e-12 if prec == "f64" else 1e-5
    for name, got, want in (("C", C.detach(), Cr.detach()), ("dA", A.grad, Ar.grad), ("dB", B.grad, Br.grad)):
        err = rel_frobenius(got.cpu().double().numpy(), want.numpy())
        assert err <= tol, (name, err)


@pytest.mark.parametrize("mode", ["persistent", "grouped"])
@pytest.mark.parametrize("batch", range(6))
def test_random_batch(batch, mode):
    """Batches of eight random cases (different patterns and extents in one
    launch) with per-problem alpha and beta: the persistent gather launch and
    the grouped packed launch (records of the same batches checked on the CPU
    by tests/test_group.py)."""
    import torch
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.batch import GettBatch
    from tccg_amd.plan import PACKED_TILES
    cs = [random_contraction(8 * batch + i) for i in range(8)]
    rng = np.random.default_rng(batch)
    alphas = [float(x) for x in rng.uniform(-2, 2, len(cs))]
    betas = [0.0 if i % 2 else float(x) for i, x in enumerate(rng.uniform(-1, 1, len(cs)))]
    dev = torch.device("cuda", 0)
    host, As, Bs, Cs = [], [], [], []
    for c in cs:
        A, B, C = (rng.standard_normal(s).astype(np.float32) for s in c.shapes())
        host.append((A, B, C))
        As.append(torch.from_numpy(A).to(dev))
        Bs.append(torch.from_numpy(B).to(dev))
        Cs.append(torch.from_numpy(C.copy()).to(dev))
    kw = dict(group_tile=PACKED_TILES[F32].index((128, 128, 16))) if mode == "grouped" else {}
    print("random batch", batch, mode, [str(c) for c in cs], flush=True)
    b = GettBatch(cs, alphas=alphas, betas=betas, **kw).bind(As, Bs, Cs)
    b.launch()
    torch.cuda.synchronize()
    for i, c in enumerate(cs):
        labels = [[x.label for x in T.indices] for T in (c.C0, c.A0, c.B0)]
        A, B, C = host[i]
        ref = contract_ref(*labels, A, B, C, alphas[i], betas[i])
        assert rel_frobenius(Cs[i].cpu().numpy(), ref) <= TOL[F32], (i, str(c))


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("seed", range(48))
def test_random_contraction_stream_k(seed, prec):
    """Stream-k plans (walk 5) of packed tile 0 on the random cases over 3, 37
    and 256 workgroups: segments inside one tile, empty segments, grids under
    the 8 XCD groups (the partition and partial-sum protocol of the same
    (grid, tiles, k-steps) triples is modelled on the CPU by
    tests/test_stream_k.py::test_random_case_partitions)."""
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.plan import build_packed_plan
    c = random_contraction(seed)
    alpha, beta = [(1.0, 0.0), (2.4, 1.1), (-0.5, 0.0)][seed % 3]
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, seed, "normal", dtype, with_c=beta != 0)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    ref = contract_ref(*labels, A, B, C, alpha, beta)
    for grid in (3, 37, 256):
        plan = build_packed_plan(c.problem, prec, 0, stream_k=1)
        plan.grid = grid
        print("random stream-k", seed, prec, grid, c, plan.describe(), flush=True)
        out = run_gpu(c, A, B, C, prec, alpha, beta, plan=plan)
        err = rel_frobenius(out, ref)
        assert np.isfinite(err) and err <= TOL[prec], (grid, plan.describe(), err)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("seed", range(48))
def test_random_contraction_split_k(seed, prec):
    """Split-k plans (k parts into partial tiles, then the ordered reduction
    writing alpha * sum + beta * C) of the 128 x 128 x 16 packed tile with 2
    and 3 parts on the random cases, wherever the plan has that many k-steps
    (the partition is modelled on the CPU by tests/test_plan_model.py)."""
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.plan import PACKED_TILES, build_packed_plan
    c = random_contraction(seed)
    alpha, beta = [(1.0, 0.0), (2.4, 1.1), (-0.5, 0.0)][seed % 3]
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, seed, "normal", dtype, with_c=beta != 0)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    ref = contract_ref(*labels, A, B, C, alpha, beta)
    tile = PACKED_TILES[prec].index((128, 128, 16))
    kt = build_packed_plan(c.problem, prec, tile).Kp // 16
    for ks in (2, 3):
        if ks > kt:
            continue
        plan = build_packed_plan(c.problem, prec, tile, ksplit=ks)
        print("random split-k", seed, prec, ks, c, plan.describe(), flush=True)
        out = run_gpu(c, A, B, C, prec, alpha, beta, plan=plan)
        err = rel_frobenius(out, ref)
        assert np.isfinite(err) and err <= TOL[prec], (ks, plan.describe(), err)


@pytest.mark.parametrize("prec", [F32, F64])
@pytest.mark.parametrize("seed", range(6))
def test_random_alpha_zero_and_accumulate(seed, prec):
    """alpha = 0 (C becomes beta * C, the products ignored) and beta = 1
    (accumulation into C) through every candidate plan of a random case."""
    from oracle.einsum_ref import contract_ref, rel_frobenius
    c = random_contraction(seed)
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, seed, "normal", dtype, with_c=True)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for alpha, beta in ((0.0, 1.5), (1.0, 1.0)):
        ref = contract_ref(*labels, A, B, C, alpha, beta)
        for i, plan in enumerate(candidate_plans(c.problem, prec)):
            print("random alpha/beta", seed, prec, alpha, beta, c, i, plan.describe(), flush=True)
            out = run_gpu(c, A, B, C, prec, alpha, beta, plan=plan)
            err = rel_frobenius(out, ref)
            assert np.isfinite(err) and err <= TOL[prec], (alpha, beta, plan.describe(), err)
