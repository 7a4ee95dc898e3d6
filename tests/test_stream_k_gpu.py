"""Stream-k packed plans (walk 5: every tile's k-steps spread evenly over the
grid, partial tiles published and added in a fixed order inside the launch)
against the FP64 einsum oracle through the C ABI: every precision, grids from
one workgroup to two per CU (segments inside one tile, empty segments, grids
under the 8 XCD groups), ragged extents, alpha/beta, the run epilogue; the
same bits from run to run and from the prepacked call; flags left clear and
the error word zero; graph replays; the generated library of a stream-k
plan; and the host checks that reject what the walk cannot run."""
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


def _sk(c, prec, tile, grid):
    plan = build_packed_plan(c.problem, prec, tile, stream_k=1)
    plan.grid = grid
    return plan


def _flags(dp):
    """The flag words and the error word of a stream-k plan's workspace."""
    import torch
    p = dp.plan
    esz = 8 if p.prec == F64 else 4
    per = {F32: 4, F64: 8, BF16X3: 6, BF16: 2, FP8: 1}[p.prec] * p.Kp
    ab = -(-(-(-(per * p.Mp) // 256) * 256 + per * p.Np) // 256) * 256
    off = ab + -(-(p.grid * p.BM * p.BN * esz) // 256) * 256
    torch.cuda.synchronize()
    return dp.workspace[off:off + 4 * (p.grid + 1)].cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3, BF16, FP8])
@pytest.mark.parametrize("grid", [1, 3, 8, 37, 256, 512])
def test_stream_k_every_precision(prec, grid):
    """Every packed tile: with grid 1 one workgroup walks every tile; 3 and
    37 leave XCD groups of unequal size; 256 and 512 cut most tiles into
    several pieces (the extents give 6 to 24 tiles of 3 to 60 k-steps)."""
    c = _long_k()
    for tile in range(len(PACKED_TILES[prec])):
        plan = _sk(c, prec, tile, grid)
        print("stream-k", grid, plan.describe(), flush=True)
        check(c, prec, seed=grid + 10 * tile, plan=plan, dist="normal", tol=TOL[prec])


RAGGED = [
    (["a", "b"], ["a", "k"], ["k", "b"]),
    (["b", "a"], ["a", "k"], ["k", "b"]),
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),
    (["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"]),
]


@pytest.mark.parametrize("pi", range(len(RAGGED)))
@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_stream_k_ragged(pi, prec):
    """Padding rows and columns in every tile and a k range that is no
    multiple of the k-step; segments shorter than a tile's k range."""
    c_l, a_l, b_l = RAGGED[pi]
    labels = sorted(set(c_l + a_l + b_l))
    ext = (5, 7, 3, 11)
    sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
    sizes["k"] = 1301 if "l" not in sizes else 263
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    for grid in (5, 64, 256):
        plan = _sk(c, prec, 0, grid)
        print("stream-k ragged", c, plan.describe(), flush=True)
        check(c, prec, seed=pi + grid, plan=plan)


@pytest.mark.parametrize("prec", [F32, F64, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (1.0, -2.0), (0.5, 0.0)])
def test_stream_k_alpha_beta(prec, alpha, beta):
    """alpha and beta applied once, by the piece that finishes each tile."""
    c = _long_k()
    check(c, prec, seed=7, alpha=alpha, beta=beta, plan=_sk(c, prec, 0, 200))


def test_stream_k_run_epilogue():
    """C[a,b,...] with a in m and b in n: the finishing pieces write C through
    the run epilogue (crun), as the one-tile walk does."""
    c = Contraction.from_labels(["a", "b", "c", "d"], ["a", "k", "c"], ["k", "d", "b"],
                                dict(a=16, b=8, c=24, d=32, k=300))
    for prec in (F32, BF16X3):
        plan = _sk(c, prec, 0, 96)
        assert plan.crun[0] > 0, plan.describe()
        print("stream-k runs", plan.describe(), flush=True)
        check(c, prec, seed=3, plan=plan, dist="normal")


def test_stream_k_bits_repeat_and_flags_clear():
    """Two calls, and a prepacked call, give the same bits (a fixed summation
    order); the call leaves every flag clear and the error word zero; the
    result differs from the one-tile walk's only by the summation order."""
    import torch
    from tccg_amd.runtime import DevicePlan
    c = _long_k()
    A, B, C = inputs(c, 3, "normal")
    one = run_gpu(c, A, B, C, F32, plan=build_packed_plan(c.problem, F32, 0))
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dp = DevicePlan(_sk(c, F32, 0, 256))
    outs = []
    for prepacked in (False, False, True):
        tC = torch.zeros(C.shape, device=dev)
        if prepacked:
            dp.run_prepacked(tC, 1.0, 0.0)
        else:
            dp.run(tA, tB, tC, 1.0, 0.0)
        fl = _flags(dp)
        assert not fl.any(), "flags or error word set after the call: %s" % np.nonzero(fl)[0][:8]
        outs.append(tC.cpu().numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    assert np.abs(outs[0] - one).max() <= 1e-5 * np.abs(one).max()


def test_stream_k_graph_replays():
    """The call captured in a HIP graph (pack kernels, flag clear, stream-k
    kernel): every replay, with eager work in between, equals the eager call."""
    import torch
    from tccg_amd.runtime import DevicePlan
    c = _long_k()
    A, B, C = inputs(c, 5, "normal")
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    dp = DevicePlan(_sk(c, F32, 0, 128))
    tC = torch.zeros(C.shape, device=dev)
    dp.run(tA, tB, tC, 1.0, 0.0)
    torch.cuda.synchronize()
    ref = tC.clone()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            dp.run(tA, tB, tC, 1.0, 0.0)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for i in range(4):
        tC.fill_(float(i + 7))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(tC, ref), "replay %d" % i
        assert not _flags(dp).any()


def test_stream_k_full_size_compute_bound():
    """abcd-aebf-dfce at full size (72/48): 729 tiles of 144 k-steps over one
    and two workgroups per CU, FP32 within the FP32 bar of the FP64 einsum."""
    from tccg_amd.workloads import get
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    for spc in (1, 2):
        plan = build_packed_plan(c.problem, F32, 3, stream_k=spc)
        print("stream-k full", plan.describe(), flush=True)
        check(c, F32, seed=spc, plan=plan, dist="normal")


def test_stream_k_rejected_before_any_launch():
    """Plans the walk cannot run fail in the host checks with C untouched:
    no grid, a split k range."""
    import torch
    from tccg_amd.runtime import DevicePlan, GettRuntimeError
    c = _long_k()
    A, B, C = inputs(c, 1)
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    tC = torch.full(C.shape, 3.0, device=dev)
    for field, value in (("grid", 0), ("ksplit", 2)):
        plan = _sk(c, F32, 0, 64)
        setattr(plan, field, value)
        with pytest.raises(GettRuntimeError):
            DevicePlan(plan).run(tA, tB, tC, 1.0, 0.0)
        torch.cuda.synchronize()
        assert bool((tC == 3.0).all()), field


def test_generated_stream_k_library():
    """The generated entry point of a stream-k plan (pack kernels, flag clear,
    stream-k launch, workspace allocated by the library) on the first adjoint
    of abcd-aklb-lckd, three calls on the same workspace."""
    import torch
    from oracle.einsum_ref import contract_ref, rel_frobenius
    from tccg_amd.runtime import GeneratedGett
    wl = "aklb-abcd-lckd"
    lib_path = codegen.GENERATED_DIR / ("lib%s_streamk1.so" % codegen.workload_library_name(wl, F32))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(codegen.SPLIT_K_TEXTS[wl])
    A, B, C = inputs(c, seed=9, dist="normal", with_c=True)
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for alpha, beta in ((1.0, 0.0), (0.5, 2.0), (1.0, 0.0)):
        tC = torch.from_numpy(C.copy()).to(dev)
        g(tA, tB, tC, alpha, beta)
        torch.cuda.synchronize()
        ref = contract_ref(*labels, A, B, C, alpha, beta)
        assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5
