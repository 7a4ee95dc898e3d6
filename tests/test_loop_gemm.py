"""LoG planning: the reference's candidate rules and selection, and a CPU
model of the grouped launch (shared GEMM tables, per-instance base offsets,
contracted loops folded into the k range) checked against the FP64 einsum
oracle, with every access asserted in bounds."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.loop_gemm import best_log_candidates, log_candidates, log_layout
from tccg_amd.plan import BF16X3, F32, F64, INVALID, PACKED_TILES
from tccg_amd.workloads import WORKLOADS, scaled


def test_known_candidates():
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    cands = log_candidates(c.problem)
    # m = a (C's stride-one), n = d (B's stride-one), k free among e, f; batch among b, c
    assert sorted(x.name for x in cands) == sorted([
        "LoG_free_c_contracted_f_batch_b_ade", "LoG_free_b_contracted_f_batch_c_ade",
        "LoG_free_c_contracted_e_batch_b_adf", "LoG_free_b_contracted_e_batch_c_adf"])
    assert all(x.gemm_size() == 72 * 72 * 48 and x.sizes[0] == 72 for x in cands)
    # A's stride-one index d is an m index other than C's: no single GEMM holds both
    assert log_candidates(Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"]).problem) is None
    # largest GEMM first, then largest batch
    t = Contraction.from_labels(["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=8, b=4, c=16, k=3))
    best = best_log_candidates(log_candidates(t.problem), 8)
    assert [x.name for x in best] == ["LoG_free_dummy_contracted_dummy_batch_b_ack"]


def emulate(c, cand, A, B, C, alpha, beta, prec=F32):
    """The grouped launch of the candidate: every instance (free loop
    combination) C[ic] = alpha * A[ia] B[ib] + beta * C[ic] over the shared
    tables, the k range folding the contracted loops."""
    pb = c.problem
    lay = log_layout(pb, cand, prec)
    BM, BN, BK = PACKED_TILES[prec][lay.tile]
    opA, opB = (B, A) if pb.swapped else (A, B)
    a, b = np.asarray(opA, np.float64).reshape(-1), np.asarray(opB, np.float64).reshape(-1)
    out = np.asarray(C, np.float64).reshape(-1).copy()
    offAm, offAk, offBn, offBk, offCm, offCn = (t.astype(np.int64) for t in lay.tables)
    # table lengths the grouped kernels assume
    assert offAm.size % BM == 0 and offAm.size % 64 == 0 and offBn.size % BN == 0 and offBn.size % 64 == 0
    assert offAk.size == offBk.size and offAk.size % BK == 0 and offAk.size % 32 == 0
    vm, vn, vk = offAm != INVALID, offBn != INVALID, offAk != INVALID
    assert np.array_equal(vm, offCm != INVALID) and np.array_equal(vn, offCn != INVALID)
    assert np.array_equal(vk, offBk != INVALID)
    Am, Ak, Bn, Bk, Cm, Cn = offAm[vm], offAk[vk], offBn[vn], offBk[vk], offCm[vm], offCn[vn]
    written = np.zeros(out.size, dtype=np.int64)
    for (fa, fb, fc) in lay.free_bases:
        ia = fa + Am[:, None] + Ak[None, :]
        ib = fb + Bk[:, None] + Bn[None, :]
        ic = fc + Cm[:, None] + Cn[None, :]
        assert ia.min() >= 0 and ia.max() < a.size and ib.min() >= 0 and ib.max() < b.size
        assert ic.min() >= 0 and ic.max() < out.size
        prod = alpha * (a[ia] @ b[ib])
        out[ic] = prod + (beta * out[ic] if beta != 0 else 0)
        written[ic.reshape(-1)] += 1
    assert np.all(written == 1)     # every C element belongs to exactly one instance
    return out.reshape(np.shape(C))


CASES = [("abcd-aebf-dfce", 0.12), ("abc-akb-kc", 0.06), ("abcd-aklb-lckd", 0.12), ("cpu48", 0.2),
         ("abcd-aebf-dfce", 0.34), ("abc-akb-kc", 0.125), ("cpu48", 0.34)]   # the last three: four-wide loads


@pytest.mark.parametrize("wl,f", CASES)
def test_model_matches_oracle(wl, f):
    c = Contraction.from_tccg(scaled(wl, f))
    rng = np.random.default_rng(7)
    sa, sb, sc = c.shapes()
    A, B, C = rng.standard_normal(sa), rng.standard_normal(sb), rng.standard_normal(sc)
    lab = lambda T: [i.label for i in T.indices]
    ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, 1.3, -0.7)
    cands = log_candidates(c.problem)
    assert cands
    for cand in cands[:6]:
        assert rel_frobenius(emulate(c, cand, A, B, C, 1.3, -0.7), ref) < 1e-13, cand.name


@pytest.mark.parametrize("prec", [F64, BF16X3])
def test_model_other_precisions(prec):
    """FP64 and split-BF16 take their own packed tile tables."""
    c = Contraction.from_tccg(scaled("abcd-aebf-dfce", 0.12))
    rng = np.random.default_rng(3)
    A, B, C = (rng.standard_normal(s) for s in c.shapes())
    lab = lambda T: [i.label for i in T.indices]
    ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, 0.5, 2.0)
    cand = log_candidates(c.problem)[0]
    assert rel_frobenius(emulate(c, cand, A, B, C, 0.5, 2.0, prec), ref) < 1e-13


def test_log_tile_and_sharing():
    """Config (i)'s GEMM (72 x 72, k = 48 x 48 folded): 128 x 128 tiles, 72
    packed A bases and 72 packed B bases shared by 5184 instances."""
    from tccg_amd.loop_gemm import log_tile
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    cand = best_log_candidates(log_candidates(c.problem), 1)[0]
    lay = log_layout(c.problem, cand)
    assert PACKED_TILES[F32][lay.tile][:2] == (128, 128) and log_tile(F32, 72, 72) == lay.tile
    assert lay.K == 48 * 48 and len(lay.free_bases) == 72 * 72
    assert len({fa for fa, _, _ in lay.free_bases}) == 72 and len({fb for _, fb, _ in lay.free_bases}) == 72
