"""Split-k packed plans (ksplit > 1: each tile's k-steps divided into ksplit
contiguous parts, one workgroup each, partial tiles added in part order by
split_reduce_kernel) against the FP64 einsum oracle through the C ABI, in
every precision, on ragged extents and with alpha/beta; the generated
library of a split-k plan; and the host checks that reject plans the split
cannot run, before any launch."""
import ctypes

import numpy as np
import pytest

from tccg_amd import Contraction, codegen
from tccg_amd.plan import BF16, BF16X3, F32, F64, FP8, PACKED_TILES, build_packed_plan

from .gpu_helpers import check, inputs, run_gpu

pytestmark = pytest.mark.gpu

TOL = {F32: None, F64: None, BF16X3: None, BF16: 1e-2, FP8: None}
# C[a,b,c] = A[a,k,l,b] * B[l,c,k]: M*N = 40*9 x 60, k = 48*20 (ragged tiles)
LONG_K = (["a", "b", "c"], ["a", "k", "l", "b"], ["l", "c", "k"], dict(a=40, b=9, c=60, k=48, l=20))


def _long_k():
    return Contraction.from_labels(*LONG_K)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3, BF16, FP8])
@pytest.mark.parametrize("ks", [2, 3, 4, 7])
def test_split_k_every_precision(prec, ks):
    c = _long_k()
    for tile in range(len(PACKED_TILES[prec])):
        plan = build_packed_plan(c.problem, prec, tile, ksplit=ks)
        if ks > plan.Kp // plan.BK:
            continue
        print("split-k", plan.describe(), flush=True)
        check(c, prec, seed=ks + 10 * tile, plan=plan, dist="normal", tol=TOL[prec])


RAGGED = [
    (["a", "b"], ["a", "k"], ["k", "b"]),
    (["b", "a"], ["a", "k"], ["k", "b"]),
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),
    (["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"]),
]


@pytest.mark.parametrize("pi", range(len(RAGGED)))
@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_split_k_ragged(pi, prec):
    """Extents with padding rows and columns in every tile, and a k range
    that is no multiple of the k-step and that the parts divide unevenly."""
    c_l, a_l, b_l = RAGGED[pi]
    labels = sorted(set(c_l + a_l + b_l))
    ext = (5, 7, 3, 11)
    sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
    sizes["k"] = 1301 if "l" not in sizes else 263    # K = 1301 or 263 * l
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    tile = 0
    for ks in (3, 5):
        plan = build_packed_plan(c.problem, prec, tile, ksplit=ks)
        assert ks <= plan.Kp // plan.BK
        print("split-k ragged", c, plan.describe(), flush=True)
        check(c, prec, seed=pi + ks, plan=plan)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (1.0, -2.0), (0.5, 0.0)])
def test_split_k_alpha_beta(prec, alpha, beta):
    c = _long_k()
    plan = build_packed_plan(c.problem, prec, 0, ksplit=3)
    check(c, prec, seed=7, alpha=alpha, beta=beta, plan=plan)


def test_split_k_is_deterministic_and_close_to_one_part():
    """The parts are added in a fixed order: two runs give the same bits; the
    split changes only the summation order against the unsplit plan."""
    c = _long_k()
    A, B, C = inputs(c, 3, "normal")
    one = run_gpu(c, A, B, C, F32, plan=build_packed_plan(c.problem, F32, 0))
    split = build_packed_plan(c.problem, F32, 0, ksplit=4)
    first = run_gpu(c, A, B, C, F32, plan=split)
    again = run_gpu(c, A, B, C, F32, plan=split)
    assert np.array_equal(first, again)
    scale = np.abs(one).max()
    assert np.abs(first - one).max() <= 1e-5 * scale


def test_split_k_rejected_before_any_launch():
    """Plans the split cannot run fail in the host checks with C untouched:
    more parts than k-steps, a persistent grid, a run epilogue."""
    import torch
    from tccg_amd.runtime import DevicePlan, GettRuntimeError
    c = _long_k()
    A, B, C = inputs(c, 1)
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    tC = torch.full(C.shape, 3.0, device=dev)
    base = build_packed_plan(c.problem, F32, 0, ksplit=2)
    for field, value in (("ksplit", base.Kp // base.BK + 1), ("grid", 1), ("crun", (4, 4))):
        plan = build_packed_plan(c.problem, F32, 0, ksplit=2)
        setattr(plan, field, value)
        with pytest.raises(GettRuntimeError):
            DevicePlan(plan).run(tA, tB, tC, 1.0, 0.0)
        torch.cuda.synchronize()
        assert bool((tC == 3.0).all()), field


def test_generated_split_k_library():
    """The generated entry point of a split-k plan (pack kernels, the (k part,
    tile) GETT launch and the reduction, workspace allocated by the library)
    on the first adjoint of abcd-aklb-lckd, twice on the same workspace."""
    import torch
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.runtime import GeneratedGett
    wl = "aklb-abcd-lckd"
    lib_path = codegen.GENERATED_DIR / ("lib%s_ksplit4.so" % codegen.workload_library_name(wl, F32))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(codegen.SPLIT_K_TEXTS[wl])
    A, B, C = inputs(c, seed=9, dist="normal", with_c=True)
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for alpha, beta in ((1.0, 0.0), (0.5, 2.0)):
        tC = torch.from_numpy(C.copy()).to(dev)
        g(tA, tB, tC, alpha, beta)
        torch.cuda.synchronize()
        ref = contract_ref(*labels, A, B, C, alpha, beta)
        assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5
