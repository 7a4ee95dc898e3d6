"""Parity of the gfx950 GETT kernels with the FP64 einsum oracle, through the
C ABI (tccg_amd.runtime -> libtccg_gfx950.so)."""
import itertools

import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.plan import BF16X3, F32, F64, FP8, FREE1, FREE4, KDIM1, KDIM4, TILES, build_plan
from tccg_amd.workloads import WORKLOADS, scaled

from .gpu_helpers import check

pytestmark = pytest.mark.gpu

SMALL = {
    "example": scaled("example", 0.5),
    "cpu48": scaled("cpu48", 0.25),
    "abcd-aebf-dfce": scaled("abcd-aebf-dfce", 0.25),
    "abc-akb-kc": scaled("abc-akb-kc", 0.125),
    "abcdef-dgfa-gceb": scaled("abcdef-dgfa-gceb", 0.34),
    "abcd-aklb-lckd": scaled("abcd-aklb-lckd", 0.25),
}


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("prec", [F32, F64, BF16X3, FP8])
def test_workloads_small_every_tile(name, prec):
    c = Contraction.from_tccg(SMALL[name])
    for tile in range(len(TILES[prec])):
        for blocked in (False, True):
            plan = build_plan(c.problem, prec, tile, m_blocked=blocked)
            check(c, prec, seed=tile, plan=plan)


# ragged extents, permutations that exercise every load mode, vector and scalar
PATTERNS = [
    (["a", "b"], ["a", "k"], ["k", "b"]),        # GEMM NN: A FREE, B KDIM
    (["a", "b"], ["k", "a"], ["k", "b"]),        # A KDIM
    (["a", "b"], ["a", "k"], ["b", "k"]),        # B FREE
    (["a", "b"], ["k", "a"], ["b", "k"]),        # both KDIM/FREE swapped
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),  # A stride-one is not C's: blocked
    (["b", "a"], ["a", "k"], ["k", "b"]),        # operands swapped internally
    (["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"]),
    (["a"], ["a", "k"], ["k"]),                  # matrix-vector (empty n set)
    (["a", "b"], ["a"], ["b"]),                  # outer product (empty k set)
]


def gather_plan(c, prec):
    """The model's best plan of the gather kernel (one workgroup per tile)."""
    from tccg_amd.plan import candidate_plans
    return next(p for p in candidate_plans(c.problem, prec) if not p.packed and p.walk == 0 and not p.grid)


@pytest.mark.parametrize("pi", range(len(PATTERNS)))
@pytest.mark.parametrize("ext", [(5, 7, 3), (8, 12, 16), (33, 17, 65)])
@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
def test_patterns_ragged(pi, ext, prec):
    """The default plan (whatever kernel the model prefers) and the gather
    kernel's plan, whose load modes these patterns cover."""
    c_l, a_l, b_l = PATTERNS[pi]
    labels = sorted(set(c_l + a_l + b_l))
    sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    check(c, prec, seed=pi)
    check(c, prec, seed=pi, plan=gather_plan(c, prec))


def test_load_modes_covered():
    """The pattern list reaches every load mode of the gather kernel for A and B."""
    seen = set()
    for c_l, a_l, b_l in PATTERNS:
        for ext in [(8, 12, 16), (5, 7, 3)]:
            labels = sorted(set(c_l + a_l + b_l))
            sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
            c = Contraction.from_labels(c_l, a_l, b_l, sizes)
            p = gather_plan(c, F32)
            seen.add(("A", p.modeA))
            seen.add(("B", p.modeB))
    for op in "AB":
        for m in (FREE4, FREE1, KDIM4, KDIM1):
            assert (op, m) in seen, (op, m, seen)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.4, 1.1), (-0.5, 1.0), (1.0, -2.0)])
def test_alpha_beta(prec, alpha, beta):
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.2))
    check(c, prec, seed=3, alpha=alpha, beta=beta)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (0.5, 0.0)])
def test_gather_persistent_alpha_beta(prec, alpha, beta):
    """The gather kernel's persistent walk (a grid below the tile count;
    write-back C stores): beta != 0 and beta == 0, ragged edge tiles."""
    c = Contraction.from_labels(["a", "b", "c"], ["a", "k", "c"], ["k", "b"], {"a": 37, "b": 70, "c": 9, "k": 45})
    plan = gather_plan(c, prec)
    plan.grid = min(3, plan.ntiles() - 1)
    assert plan.grid > 0
    print("gather persistent", prec, alpha, beta, plan.describe(), flush=True)
    check(c, prec, seed=5, alpha=alpha, beta=beta, plan=plan)


def test_beta_from_tccg():
    txt = """C[i1,j,i2] = 2.4 * A[i1,k,i2] * B[j,k] + 1.1 * C[i1,j,i2]
i1 = 48
i2 = 40
j = 36
k = 8
"""
    c = Contraction.from_tccg(txt)
    assert (c.alpha, c.beta) == (2.4, 1.1)
    check(c, F32, seed=5)


def test_fused_indices_run():
    # C[a,b,c] = A[a,b,k] * B[k,c] fuses (a,b) -> one index
    c = Contraction.from_labels(["a", "b", "c"], ["a", "b", "k"], ["k", "c"], dict(a=12, b=10, c=20, k=9))
    assert len(c.C.indices) == 2
    check(c, F32, seed=7)


def _golden_cases():
    import glob
    import os
    return sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", _golden_cases(), ids=lambda p: p.split("/")[-1])
@pytest.mark.parametrize("prec", [F32, BF16X3])
def test_golden_fixtures(path, prec):
    """GPU result against the committed FP64 fixture (inputs stored as FP32)."""
    import torch
    from oracle.einsum_ref import rel_frobenius
    z = np.load(path, allow_pickle=False)
    sizes = dict(zip([str(x) for x in z["size_labels"]], [int(x) for x in z["size_values"]]))
    c = Contraction.from_labels([str(x) for x in z["c"]], [str(x) for x in z["a"]], [str(x) for x in z["b"]],
                                sizes, float(z["alpha"]), float(z["beta"]))
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(z[k])).to(dev) for k in ("A", "B", "C_in"))
    c(tA, tB, tC, prec=prec)
    torch.cuda.synchronize()
    assert rel_frobenius(tC.cpu().numpy(), z["C_out"]) < 1e-6


# single-k-step contractions (K <= 8): the streaming walk, every tile that has it
STREAM_CASES = [
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=96, b=40, c=72, k=8)),   # bandwidth-bound pattern
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=36, b=17, c=65, k=4)),   # ragged tiles, k padded
    (["a", "b"], ["a", "k"], ["b", "k"], dict(a=200, b=96, k=8)),                 # B along k
]


@pytest.mark.parametrize("ci", range(len(STREAM_CASES)))
@pytest.mark.parametrize("grid", [None, 3])
def test_stream_walk(ci, grid):
    from tccg_amd.plan import stream_ok
    c_l, a_l, b_l, sizes = STREAM_CASES[ci]
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    ran = 0
    for tile, (BM, BN, BK) in enumerate(TILES[F32]):
        if BK != 8:
            continue
        plan = build_plan(c.problem, F32, tile, stream=True)
        if plan.walk != 2:
            assert not stream_ok(plan)
            continue
        if grid is not None:
            plan.grid = min(grid, plan.ntiles())
        print("stream", ci, plan.describe(), flush=True)
        check(c, F32, seed=ci, plan=plan)
        check(c, F32, seed=ci, plan=plan, alpha=2.0, beta=-0.5)
        ran += 1
    assert ran > 0
