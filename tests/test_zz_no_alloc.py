"""The no-allocation packed entry point (tccg_gett_packed_noalloc) through the
C ABI, against the packed entry point with a caller workspace
(tccg_gett_packed_<prec>): same kernels, so the same bits.  Two calls in a
row on the shared device array, the second with a smaller contraction, and
a plan beyond the array's size rejected before any launch."""
import ctypes

import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.plan import BF16X3, F32, FP8, PACKED_TILES, build_packed_plan
from tccg_amd.workloads import WORKLOADS, scaled

from .gpu_helpers import inputs

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20


def _noalloc(dp, tA, tB, tC, alpha, beta):
    """Run dp's plan through tccg_gett_packed_noalloc (FP8: its amax first,
    as DevicePlan.run does)."""
    import torch
    from tccg_amd.runtime import load_library, last_error
    lib = load_library()
    pb = dp.plan.problem
    opA, opB = (tB, tA) if pb.swapped else (tA, tB)
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    if dp.plan.prec == FP8:
        assert lib.tccg_amax_f32(opA.data_ptr(), opA.numel(), dp.amax.data_ptr(), sp) == 0
        assert lib.tccg_amax_f32(opB.data_ptr(), opB.numel(), dp.amax.data_ptr() + 4, sp) == 0
    rc = lib.tccg_gett_packed_noalloc(dp.plan.prec, opA.data_ptr(), opB.data_ptr(), tC.data_ptr(), alpha, beta,
                                      ctypes.byref(dp.cplan), sp)
    assert rc == 0, last_error()


@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_noalloc_matches_packed_entry_point(prec):
    import torch
    from tccg_amd.runtime import DevicePlan
    dev = torch.device("cuda", 0)
    cases = [(Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"]), 1.0, 0.0),          # first call
             (Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.25)), 1.5, 0.5)]       # smaller, second
    for i, (c, alpha, beta) in enumerate(cases):
        tile = PACKED_TILES[prec].index((128, 128, 128) if prec == FP8 else (128, 128, 16))
        plan = build_packed_plan(c.problem, prec, tile, persistent=(i == 0))
        assert plan.workspace_bytes() <= SCRATCH_BYTES
        if i:
            assert plan.workspace_bytes() < prev
        prev = plan.workspace_bytes()
        dp = DevicePlan(plan)
        A, B, C = inputs(c, 20 + i, "normal", with_c=True)
        tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
        want = torch.from_numpy(C).to(dev)
        got = want.clone()
        print("no-alloc call %d: %s %s" % (i, c, plan.describe()), flush=True)
        dp.run(tA, tB, want, alpha, beta)
        _noalloc(dp, tA, tB, got, alpha, beta)
        torch.cuda.synchronize()
        assert torch.equal(got, want), "call %d differs from tccg_gett_packed_<prec>" % i


def test_noalloc_rejects_plans_beyond_the_scratch():
    import torch
    from tccg_amd.runtime import DevicePlan, load_library
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    plan = build_packed_plan(c.problem, BF16X3, PACKED_TILES[BF16X3].index((128, 128, 16)))
    assert plan.workspace_bytes() > SCRATCH_BYTES
    dp = DevicePlan(plan)
    lib = load_library()
    rc = lib.tccg_gett_packed_noalloc(BF16X3, None, None, None, 1.0, 0.0, ctypes.byref(dp.cplan),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc != 0 and b"no-allocation" in lib.tccg_last_error()
