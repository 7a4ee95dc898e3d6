"""Parity of the vector micro-kernel (walk 3, contractions with K <= 16) with
the FP64 einsum oracle, through the C ABI (tccg_gett_f32 / tccg_gett_f64).
Every case here is one whose accesses tests/test_vector_plan.py models on the
CPU."""
import pytest

from tccg_amd.plan import F32, F64, VECTOR_TILES, build_vector_plan

from .gpu_helpers import check
from .test_vector_plan import CASES, _contraction

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("prec", [F32, F64])
def test_vector_every_tile(ci, prec):
    c = _contraction(ci)
    ran = 0
    for tile, (BM, BN, BK) in enumerate(VECTOR_TILES[prec]):
        if c.problem.K > BK:
            continue
        plan = build_vector_plan(c.problem, prec, tile)
        print("vector", ci, plan.describe(), flush=True)
        check(c, prec, seed=ci, plan=plan)
        check(c, prec, seed=ci + 100, plan=plan, alpha=2.0, beta=-0.5)
        ran += 1
    assert ran > 0


def test_vector_rejects_other_precisions():
    import ctypes

    import torch

    from tccg_amd.plan import BF16X3
    from tccg_amd.runtime import DevicePlan, load_library
    c = _contraction(0)
    dp = DevicePlan(build_vector_plan(c.problem, F32, 0))
    lib = load_library()
    A, B, C = (torch.zeros(s, dtype=torch.float32, device="cuda") for s in c.shapes())
    rc = lib.tccg_gett_bf16x3(A.data_ptr(), B.data_ptr(), C.data_ptr(), 1.0, 0.0, ctypes.byref(dp.cplan), None)
    assert rc != 0 and "FP32 or FP64" in lib.tccg_last_error().decode()
    assert BF16X3 not in VECTOR_TILES
