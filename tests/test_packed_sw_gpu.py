"""Parity of the store-wave packed kernel (walk 4: compute waves and store
waves in one workgroup, C staged through LDS) with the FP64 einsum oracle
and bit for bit with the packed kernel of the same tile, through the C ABI."""
import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.plan import BF16, BF16X3, F32, FP8, PACKED_TILES, build_packed_plan, sw_built
from tccg_amd.workloads import scaled

from .gpu_helpers import check, inputs, run_gpu

pytestmark = pytest.mark.gpu

SW_TILES = [(prec, t) for prec in (F32, BF16X3, BF16, FP8) for t in range(len(PACKED_TILES[prec]))
            if sw_built(prec, t)]
TOL = {F32: None, BF16X3: None, BF16: 1e-2, FP8: None}

SMALL = {
    "abcd-aebf-dfce": scaled("abcd-aebf-dfce", 0.25),
    "abcdef-dgfa-gceb": scaled("abcdef-dgfa-gceb", 0.34),
    "abc-akb-kc": scaled("abc-akb-kc", 0.125),
}


def test_store_wave_tiles_exist():
    assert {p for p, _ in SW_TILES} == {F32, BF16X3, BF16, FP8}


@pytest.mark.parametrize("prec,tile", SW_TILES)
@pytest.mark.parametrize("name", sorted(SMALL))
def test_store_waves_every_tile(prec, tile, name):
    c = Contraction.from_tccg(SMALL[name])
    plan = build_packed_plan(c.problem, prec, tile, store_waves=True)
    print("store waves", name, plan.describe(), flush=True)
    check(c, prec, seed=tile, plan=plan, tol=TOL[prec])


@pytest.mark.parametrize("prec,tile", SW_TILES)
@pytest.mark.parametrize("grid", [1, 3, 7])
def test_store_waves_walk_many_tiles(prec, tile, grid):
    """Small grids: every workgroup walks many tiles (single- and multi-step
    k ranges), so the staged tile and the offset buffers are reused."""
    for k in (8, 96):
        c = Contraction.from_labels(["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=40, b=9, c=150, k=k))
        plan = build_packed_plan(c.problem, prec, tile, store_waves=True)
        plan.grid = grid
        print("store waves walk", c, plan.describe(), flush=True)
        check(c, prec, seed=grid, plan=plan, tol=TOL[prec])


RAGGED = [
    (["a", "b"], ["a", "k"], ["k", "b"]),
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),
    (["b", "a"], ["a", "k"], ["k", "b"]),
    (["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"]),
]


@pytest.mark.parametrize("pi", range(len(RAGGED)))
@pytest.mark.parametrize("ext", [(5, 7, 3), (33, 17, 65)])
@pytest.mark.parametrize("prec", [F32, BF16X3, FP8])
def test_store_waves_ragged(pi, ext, prec):
    """Extents that are no multiple of 4 leave C without 16-byte groups
    (cvec 0: element stores), and tiles with padding rows and columns."""
    c_l, a_l, b_l = RAGGED[pi]
    labels = sorted(set(c_l + a_l + b_l))
    sizes = {l: ext[i % len(ext)] for i, l in enumerate(labels)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    tile = next(t for p, t in SW_TILES if p == prec)
    plan = build_packed_plan(c.problem, prec, tile, store_waves=True)
    check(c, prec, seed=pi, plan=plan)


@pytest.mark.parametrize("prec", [F32, BF16X3])
@pytest.mark.parametrize("alpha,beta", [(2.4, 1.1), (1.0, -2.0)])
def test_store_waves_alpha_beta(prec, alpha, beta):
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.2))
    tile = next(t for p, t in SW_TILES if p == prec)
    check(c, prec, seed=3, alpha=alpha, beta=beta, plan=build_packed_plan(c.problem, prec, tile, store_waves=True))


@pytest.mark.parametrize("prec,tile", SW_TILES)
def test_store_waves_bitwise_equal_packed(prec, tile):
    """Same tile, same k order, same conversion: the store-wave kernel gives
    the packed kernel's bits (at alpha = 1 the FP8 factor alpha * descale is
    the packed kernel's descale)."""
    c = Contraction.from_tccg(SMALL["abcdef-dgfa-gceb"])
    A, B, C = inputs(c, 5, "normal")
    sw = build_packed_plan(c.problem, prec, tile, store_waves=True)
    ref = build_packed_plan(c.problem, prec, tile)
    got = run_gpu(c, A, B, C, prec, plan=sw)
    want = run_gpu(c, A, B, C, prec, plan=ref)
    assert np.array_equal(got, want)
