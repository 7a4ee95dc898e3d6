"""Parity of the packed GETT path (operands converted once into tile-major
planes, then copied tile by tile) with the FP64 einsum oracle and with the
unpacked kernels, through the C ABI (tccg_gett_packed_<prec>)."""
import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.plan import BF16, BF16X3, F32, F64, FP8, PACKED_TILES, TILES, build_packed_plan, build_plan
from tccg_amd.workloads import scaled

from .gpu_helpers import check, inputs, run_gpu

pytestmark = pytest.mark.gpu

TOL_BF16 = 1e-2   # one BF16 product per term: ~2^-9 relative per operand

SMALL = {
    "example": scaled("example", 0.5),
    "abcd-aebf-dfce": scaled("abcd-aebf-dfce", 0.25),
    "abc-akb-kc": scaled("abc-akb-kc", 0.125),
    "abcdef-dgfa-gceb": scaled("abcdef-dgfa-gceb", 0.34),
    "abcd-aklb-lckd": scaled("abcd-aklb-lckd", 0.25),
}


def _tol(prec):
    return TOL_BF16 if prec == BF16 else None


@pytest.mark.parametrize("name", sorted(SMALL))
@pytest.mark.parametrize("prec", [F32, F64, BF16X3, BF16, FP8])
def test_packed_every_tile(name, prec):
    c = Contraction.from_tccg(SMALL[name])
    for tile in range(len(PACKED_TILES[prec])):
        plan = build_packed_plan(c.problem, prec, tile)
        print("packed", name, prec, plan.describe(), flush=True)
        check(c, prec, seed=tile, plan=plan, tol=_tol(prec))


RAGGED = [
    (["a", "b"], ["a", "k"], ["k", "b"]),
    (["a", "b"], ["k", "a"], ["b", "k"]),
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),
    (["b", "a"], ["a", "k"], ["k", "b"]),
    (["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"]),
    (["a"], ["a", "k"], ["k"]),
    (["a", "b"], ["a"], ["b"]),
]


@pytest.mark.parametrize("pi", range(len(RAGGED)))
@pytest.mark.parametrize("ext", [(5, 7, 3), (33, 17, 65)])
@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
def test_packed_ragged(pi, ext, prec):
    c_l, a_l, b_l = RAGGED[pi]
    labels = sorted(set(c_l + a_l + b_l))
    sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    plan = build_packed_plan(c.problem, prec, 0)
    check(c, prec, seed=pi, plan=plan)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (1.0, -2.0)])
def test_packed_alpha_beta(prec, alpha, beta):
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.2))
    check(c, prec, seed=3, alpha=alpha, beta=beta, plan=build_packed_plan(c.problem, prec, 0))


@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_packed_persistent_walk(prec):
    """A grid smaller than the tile count: each workgroup walks several tiles."""
    c = Contraction.from_tccg(scaled("abcdef-dgfa-gceb", 0.34))
    plan = build_packed_plan(c.problem, prec, 0)
    assert plan.ntiles() > 8
    plan.grid = 7
    check(c, prec, seed=11, plan=plan)


@pytest.mark.parametrize("steps", [1, 2, 3, 5])
@pytest.mark.parametrize("prec", [F32, F64, BF16X3, FP8])
def test_packed_persistent_every_tile(prec, steps):
    """The persistent walk of every packed tile
    with a few k-steps per tile (`steps` pack blocks of k; 16-deep tiles get
    twice as many): tiles shorter than the ring, where the waits after one
    epilogue run into the next, and longer ones."""
    for tile, (BM, BN, BK) in enumerate(PACKED_TILES[prec]):
        kb = max(BK, 32)    # k is padded to whole 32-deep pack blocks
        c = Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"],
                                    {"a": 2 * BM + 17, "b": 3 * BN + 5, "k": steps * kb - 3})
        plan = build_packed_plan(c.problem, prec, tile)
        assert plan.ntiles() > 5 and plan.Kp == steps * kb, (plan.ntiles(), plan.Kp)
        plan.grid = 5
        print("persistent", prec, steps, plan.describe(), flush=True)
        check(c, prec, seed=tile + 10 * steps, plan=plan, tol=_tol(prec))


@pytest.mark.parametrize("prec", [F32, F64, BF16X3, FP8])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (1.0, -2.0), (0.5, 0.0)])
def test_packed_persistent_alpha_beta(prec, alpha, beta):
    """The persistent walk's epilogue (write-back C stores): C read and
    updated with beta != 0, and with beta == 0 both the whole tiles (no
    per-element checks) and the ragged edge tiles."""
    BM, BN, _ = PACKED_TILES[prec][0]
    c = Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], {"a": 2 * BM + 17, "b": 3 * BN + 5, "k": 61})
    plan = build_packed_plan(c.problem, prec, 0)
    plan.grid = 5
    assert plan.ntiles() > plan.grid
    print("persistent alpha/beta", prec, alpha, beta, plan.describe(), flush=True)
    check(c, prec, seed=21, alpha=alpha, beta=beta, plan=plan, tol=_tol(prec))


# (precision, packed tile id, unpacked tile id) with the same (BM, BN, BK)
SAME_SHAPE = [(F32, 0, 0), (F32, 2, 5), (F64, 0, 0), (BF16X3, 1, 0), (BF16X3, 0, 3), (BF16, 0, 0), (FP8, 0, 0)]


@pytest.mark.parametrize("prec,tp,tu", SAME_SHAPE)
def test_packed_bitwise_equals_unpacked(prec, tp, tu):
    """Same tile shape, same k order, same conversion arithmetic: the packed
    and unpacked kernels produce identical bits."""
    assert PACKED_TILES[prec][tp] == TILES[prec][tu]
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.25))
    pp = build_packed_plan(c.problem, prec, tp)
    pu = build_plan(c.problem, prec, tu, m_blocked=False)
    # same k order; the packed plan may append whole k-steps of padding (zeros)
    assert list(pp.offAk[:pu.Kp]) == list(pu.offAk)
    assert np.all(pp.offAk[pu.Kp:] == np.uint32(0xFFFFFFFF))
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, 21, "normal", dtype)
    out_p = run_gpu(c, A, B, C, prec, plan=pp)
    out_u = run_gpu(c, A, B, C, prec, plan=pu)
    assert np.array_equal(out_p.view(np.uint8), out_u.view(np.uint8))


def test_packed_workspace_too_small_is_rejected():
    import ctypes

    import torch

    from tccg_amd.runtime import DevicePlan, load_library
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.2))
    dp = DevicePlan(build_packed_plan(c.problem, BF16X3, 0))
    lib = load_library()
    A, B, C = (torch.zeros(s, dtype=torch.float32, device="cuda") for s in c.shapes())
    opA, opB = (B, A) if c.problem.swapped else (A, B)
    rc = lib.tccg_gett_packed_bf16x3(opA.data_ptr(), opB.data_ptr(), C.data_ptr(), 1.0, 0.0,
                                     ctypes.byref(dp.cplan), dp.workspace.data_ptr(), dp.workspace.numel() - 1,
                                     None)
    assert rc != 0
    assert "workspace" in lib.tccg_last_error().decode()


@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.4, 1.1)])
def test_packed_run_epilogue(prec, alpha, beta):
    """C's contiguous runs span m and n (the precision contraction's layout at
    extents 12): the 192 x 192 tile writes whole runs staged through LDS."""
    text = """C[a,b,c,d,e,f] = A[d,g,f,a] * B[g,c,e,b]
a = 12
b = 12
c = 12
d = 12
e = 12
f = 12
g = 40
"""
    c = Contraction.from_tccg(text)
    tile = PACKED_TILES[prec].index(next(t for t in PACKED_TILES[prec] if t[:2] == (192, 192)))
    plan = build_packed_plan(c.problem, prec, tile)
    assert plan.crun[0] == 12 and plan.crun[1] > 12, plan.crun
    check(c, prec, seed=13, alpha=alpha, beta=beta, plan=plan, tol=(8e-2 if prec == FP8 else None))


@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_prepacked_reuses_packed_operands(prec):
    """tccg_gett_prepacked over the operands the previous call packed gives
    the same bits as the full call."""
    import torch
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.25))
    dp = c.device_plan(plan=build_packed_plan(c.problem, prec, 0))
    A, B, C = inputs(c, 2, "normal")
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (A, B, C))
    dp.run(tA, tB, tC)
    first = tC.clone()
    tC.zero_()
    dp.run_prepacked(tC)
    torch.cuda.synchronize()
    assert torch.equal(first, tC)


def test_bad_tensors_are_rejected_before_launch():
    """Host tensors, wrong element counts and wrong dtypes raise before any
    kernel is launched, for run() and run_prepacked() alike."""
    import torch

    from tccg_amd.runtime import GettRuntimeError
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.2))
    dp = c.device_plan(plan=build_packed_plan(c.problem, F32, 0))
    A, B, C = (torch.zeros(s, dtype=torch.float32, device="cuda") for s in c.shapes())
    with pytest.raises(GettRuntimeError, match="tables are on"):
        dp.run(A.cpu(), B, C)
    with pytest.raises(GettRuntimeError, match="elements"):
        dp.run(A, B, C.reshape(-1)[:-1])
    with pytest.raises(TypeError, match="dtype"):
        dp.run(A, B.double(), C)
    dp.run(A, B, C)
    torch.cuda.synchronize()
    with pytest.raises(GettRuntimeError, match="elements"):
        dp.run_prepacked(C.reshape(-1)[:-1])
