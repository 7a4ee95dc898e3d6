"""The restated AVX2 planner (oracle/avx2_model.py): the reference's
performance model (tccg/gett.py:590-687, tccg/arch.py:21-55) pinned by
estimates worked out by hand for the best-ranked candidate of each CPU
configuration, and the candidate search's rules (tccg/gett.py:961-1150)."""
import pytest

from oracle.avx2_model import (Avx2Arch, Avx2Gett, VARIANTS, candidates_for, enumerate_candidates,
                               get_possible_values, rank_candidates)
from tccg_amd.ir import Index, Tensor

BW = 13.0 * 2 ** 30 * 0.97          # axpy bandwidth x (1 - 2 * 0.015), B/s
PEAK = 2.5 * 2 * 8 * 2              # 80 GFLOP/s single precision

CFG1 = (["a", "b", "c", "d"], ["a", "k", "c"], ["k", "d", "b"], dict(a=48, b=48, c=48, d=48, k=64))
CFG_I = (["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"], dict(a=72, b=72, c=72, d=72, e=48, f=48))
CFG_II = (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=384, b=384, c=384, k=8))


def est(t_gemm, packing):
    return PEAK * t_gemm / (packing + t_gemm)


def test_arch_constants():
    s, d = Avx2Arch("float"), Avx2Arch("double")
    assert (s.peak_gflops(), d.peak_gflops()) == (80.0, 40.0)
    assert (s.register_size, s.cache_line_size, d.register_size, d.cache_line_size) == (8, 16, 4, 8)


def test_config_1_best_candidate_by_hand():
    """var0 mc=192 nc=2304 kc=64: loops n(1) k(1) m(12).  B packed inside n,k:
    2304*64*4 B; A inside n,k,m: 192*64*4*12 B; C: 192*2304*4*12 B.  B~'s first
    index is an nr=4 sub-index, not B^'s first (k): factor 0.8 and vector
    efficiency 0.95.  Caches: A~ + 2 B slivers = (192*64 + 2*4*64)*4 B fits
    7/8 of L2; L1 (2*4*64 + 24*64)*4 B fits: efficiency 0.95."""
    packing = (2 * 2304 * 64 * 4 / (BW * 0.8 * 0.95) + 2 * 192 * 64 * 4 * 12 / BW
               + 2 * 192 * 2304 * 4 * 12 / BW)
    t_gemm = 2304 * 2304 * 64 * 2 / 1e9 / (PEAK * 0.95)
    want = est(t_gemm, packing)
    assert abs(want - 58.2503) < 1e-3
    top = candidates_for(*CFG1)[0]
    assert (top.variant, top.mc, top.nc, top.kc) in ((0, 192, 2304, 64), (1, 2304, 192, 64))
    assert top.estimate == pytest.approx(want, rel=1e-12)


def test_config_ii_best_candidate_by_hand():
    """var0 mc=192 nc=384 kc=8: loops n(1) k(1) m(768).  A^ = a_0_0(24) a_0_1(8)
    k(8) is contiguous for 192 elements, B^ for 3072, C^ for 192 (all whole
    cache lines); B~ starts with c_0 (extent 4 < 8) at position 1 of B^:
    0.8 and 0.95; no cache penalty."""
    packing = (2 * 384 * 8 * 4 / (BW * 0.8 * 0.95) + 2 * 192 * 8 * 4 * 768 / BW + 2 * 192 * 384 * 4 * 768 / BW)
    t_gemm = 147456 * 384 * 8 * 2 / 1e9 / (PEAK * 0.95)
    want = est(t_gemm, packing)
    assert abs(want - 20.698) < 1e-3
    top = candidates_for(*CFG_II)[0]
    assert (top.variant, top.mc, top.nc, top.kc) == (0, 192, 384, 8)
    assert top.estimate == pytest.approx(want, rel=1e-12)


def test_config_i_best_candidate_by_hand():
    """var1 mc=2592 nc=192 kc=256: loops m(2) k(9) n(27).  A packed inside m,k:
    2592*256*4*18 B, contiguous run 1152 (whole lines); B inside m,k,n:
    192*256*4*486 B, contiguous run 24 elements: line efficiency 24/32, its
    first index d_0_0 has extent 4: 0.95; C: 2592*192*4*486 B.  L2:
    (192*256 + 2*24*256)*4 B = 245760 > 7/8 * 256 KiB: 0.85; L1:
    (4*256 + 2*24*256)*4 B = 53248 > 7/8 * 32 KiB: 0.85."""
    packing = (2 * 2592 * 256 * 4 * 18 / BW + 2 * 192 * 256 * 4 * 486 / (BW * 0.75 * 0.95)
               + 2 * 2592 * 192 * 4 * 486 / BW)
    t_gemm = 5184 * 5184 * 2304 * 2 / 1e9 / (PEAK * 0.95 * 0.85 * 0.85)
    want = est(t_gemm, packing)
    assert abs(want - 74.3993) < 1e-3
    top = candidates_for(*CFG_I)[0]
    assert (top.variant, top.mc, top.nc, top.kc) == (1, 2592, 192, 256)
    assert top.estimate == pytest.approx(want, rel=1e-12)
    assert top.name().startswith("gett_var1_mc2592_nc192_kc256_mc124_nc14_mr24_nr4_")


def test_beta_packs_c_three_times():
    """With beta != 0 C's traffic counts three times instead of twice."""
    c, a, b, sizes = CFG1
    mk = lambda lab, idx: Tensor(lab, [Index(l, sizes[l]) for l in idx])
    g0 = Avx2Gett(mk("A", a), mk("B", b), mk("C", c))
    g1 = Avx2Gett(mk("A", a), mk("B", b), mk("C", c), beta=1.0)
    e0 = max(x.estimate for x in enumerate_candidates(g0))
    e1 = max(x.estimate for x in enumerate_candidates(g1))
    assert e1 < e0


@pytest.mark.parametrize("cfg", [CFG1, CFG_I, CFG_II])
def test_candidate_search_rules(cfg):
    c, a, b, sizes = cfg
    mk = lambda lab, idx: Tensor(lab, [Index(l, sizes[l]) for l in idx])
    g = Avx2Gett(mk("A", a), mk("B", b), mk("C", c))
    every = enumerate_candidates(g)
    ranked = rank_candidates(every, -1)
    assert len({x.key() for x in ranked}) == len(ranked)
    assert all(x.estimate >= y.estimate for x, y in zip(ranked, ranked[1:]))
    for x in ranked:
        assert g.sizeM % x.mc == 0 and g.sizeN % x.nc == 0 and g.sizeK % x.kc == 0
        assert x.mc % x.mr == 0 and x.nc % x.nr == 0 and (x.mr, x.nr, x.mc1, x.nc1) == (24, 4, 24, 4)
        # the n-outer variant packs all of N (below 4600) per k-block, the m-outer one all of M
        if x.variant == 0 and g.sizeN < 4600:
            assert x.nc == g.sizeN
        if x.variant == 1 and g.sizeM < 4600:
            assert x.mc == g.sizeM
        # each key keeps its best-rated split
        assert x.estimate == max(y.estimate for y in every if y.key() == x.key())
        # only the largest macro-kernel among equal estimates of one variant
        assert all(y.mc * y.nc * y.kc <= x.mc * x.nc * x.kc for y in ranked
                   if y.variant == x.variant and y.estimate == x.estimate)
    assert len(rank_candidates(every, 16)) == min(16, len(ranked))
    assert {x.variant for x in ranked} == {0, 1}
    assert VARIANTS[0][0] == "n" and VARIANTS[1][0] == "m"
