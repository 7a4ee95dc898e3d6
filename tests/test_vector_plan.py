"""Vector micro-kernel plans (contractions with K <= 16): the table contract
the kernel relies on, and a CPU model of the kernel's addressing
(csrc/gett_vector.hpp) whose result is checked against the FP64 einsum oracle.
The model asserts that every table index and element access the kernel makes
stays in bounds, so the GPU runs only launches whose accesses were shown to
be in bounds here."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction, codegen
from tccg_amd.plan import (BF16X3, F32, F64, FREE1, FREE4, INVALID, KIND_VECTOR, VECTOR_MAX_K, VECTOR_TILES,
                           build_vector_plan, candidate_plans, plan_from_kind, plan_kind, vector_ok)

# (C, A, B, extents, vector groups expected)
CASES = [
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=96, b=40, c=72, k=8), True),    # bandwidth-bound pattern
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=36, b=17, c=65, k=4), True),    # ragged, m and n padded
    (["a", "b"], ["a", "k"], ["b", "k"], dict(a=200, b=96, k=8), True),                  # B along k
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=5, b=7, c=3, k=3), False),      # odd extents: scalar
    (["a", "b"], ["k", "a"], ["k", "b"], dict(a=64, b=33, k=16), False),                 # A along k: scalar
    (["a", "b"], ["a", "k"], ["k", "b"], dict(a=128, b=9, k=1), True),                   # K = 1
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"], dict(a=12, b=20, c=8, k=6), False),    # C's index not A's
]


def _contraction(ci):
    c_l, a_l, b_l, sizes, _ = CASES[ci]
    return Contraction.from_labels(c_l, a_l, b_l, sizes)


def emulate(plan, A, B, C, alpha, beta):
    """What gett_vector_kernel computes, by its own addressing: groups of VEC
    m positions starting at multiples of VEC, A read at offAm[m0] + offAk[k] + v,
    C written at offCm[m0] + offCn[n] + v."""
    VEC = 4 if plan.modeA == FREE4 else 1
    assert plan.walk == 3 and plan.BM % (256 * VEC) == 0
    assert plan.Mp % plan.BM == 0 and plan.Np % plan.BN == 0 and plan.Kp == plan.BK
    a, b, c = A.reshape(-1), B.reshape(-1), C.reshape(-1).copy()
    m0 = np.arange(0, plan.Mp, VEC)
    oa = plan.offAm[m0].astype(np.int64)
    cm = plan.offCm[m0].astype(np.int64)
    ok_a = plan.offAk.astype(np.int64)
    ok_b = plan.offBk.astype(np.int64)
    ob = plan.offBn.astype(np.int64)
    cn = plan.offCn.astype(np.int64)
    inv = np.int64(INVALID)
    gv = oa != inv
    assert np.array_equal(gv, cm != inv)
    if VEC == 4:
        # the FREE4 contract: whole groups, contiguous and 16-byte aligned in A and C
        for v in range(4):
            assert np.array_equal(plan.offAm[m0 + v].astype(np.int64)[gv], oa[gv] + v)
            assert np.array_equal(plan.offCm[m0 + v].astype(np.int64)[gv], cm[gv] + v)
            assert np.all(plan.offAm[m0 + v][~gv] == INVALID)
        kv = ok_a != inv
        assert np.all((oa[gv][:, None] + ok_a[kv][None, :]) % 4 == 0)
        assert np.all(cm[gv] % 4 == 0) and np.all(cn[cn != inv] % 4 == 0)
    kva, kvb, nv = ok_a != inv, ok_b != inv, (ob != inv) & (cn != inv)
    assert np.array_equal(ob != inv, cn != inv)
    # element accesses in bounds
    if gv.any() and kva.any():
        assert (oa[gv].max() + ok_a[kva].max() + VEC - 1) < a.size
    if nv.any() and kvb.any():
        assert ob[nv].max() + ok_b[kvb].max() < b.size
    if gv.any() and nv.any():
        assert cm[gv].max() + cn[nv].max() + VEC - 1 < c.size
    # B block as the workgroups stage it (zero where invalid), A groups as held in registers
    Bm = np.zeros((plan.Np, plan.BK))
    vb = nv[:, None] & kvb[None, :]
    Bm[vb] = b[(ob[:, None] + ok_b[None, :])[vb]]
    for v in range(VEC):
        Am = np.zeros((m0.size, plan.BK))
        va = gv[:, None] & kva[None, :]
        Am[va] = a[(oa[:, None] + ok_a[None, :] + v)[va]]
        out = alpha * (Am[gv] @ Bm[nv].T)
        idx = cm[gv][:, None] + v + cn[nv][None, :]
        c[idx] = out + (beta * c[idx] if beta != 0 else 0)
    return c.reshape(C.shape)


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("prec", [F32, F64])
def test_vector_plan_model_matches_oracle(ci, prec):
    c = _contraction(ci)
    rng = np.random.default_rng(ci)
    sa, sb, sc = c.shapes()
    A, B, C = rng.uniform(-1, 1, sa), rng.uniform(-1, 1, sb), rng.uniform(-1, 1, sc)
    ref = contract_ref([i.label for i in c.C0.indices], [i.label for i in c.A0.indices],
                       [i.label for i in c.B0.indices], A, B, C, 1.5, -0.5)
    ran = 0
    for tile, (BM, BN, BK) in enumerate(VECTOR_TILES[prec]):
        if c.problem.K > BK:
            with pytest.raises(ValueError):
                build_vector_plan(c.problem, prec, tile)
            continue
        plan = build_vector_plan(c.problem, prec, tile)
        assert (plan.modeA == FREE4) == CASES[ci][4], plan.describe()
        opA, opB = (B, A) if c.problem.swapped else (A, B)
        out = emulate(plan, opA, opB, C, 1.5, -0.5)
        assert rel_frobenius(out, ref) < 1e-13
        ran += 1
    assert ran > 0


def test_vector_plan_kind_and_candidates():
    c = _contraction(0)
    plans = [p for p in candidate_plans(c.problem, F32) if p.walk == 3]
    assert len(plans) == sum(1 for t in VECTOR_TILES[F32] if t[2] >= c.problem.K)
    # the four-wide vector plan is the model's first choice for the store-bound contraction
    assert candidate_plans(c.problem, F32)[0].walk == 3
    for p in plans:
        assert plan_kind(p) == KIND_VECTOR and " vector" in p.describe()
        q = plan_from_kind(c.problem, F32, p.tile, False, plan_kind(p))
        assert q.describe() == p.describe() and np.array_equal(q.offCm, p.offCm)
    assert not any(p.walk == 3 for p in candidate_plans(c.problem, BF16X3))
    big = Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], dict(a=64, b=64, k=VECTOR_MAX_K + 1))
    assert not vector_ok(big.problem, F32)
    assert not any(p.walk == 3 for p in candidate_plans(big.problem, F32))
    with pytest.raises(ValueError):
        build_vector_plan(big.problem, F32, 0)


def test_emit_vector_source():
    text = "C[a,b,c] = A[a,k,b] * B[k,c]\na = 96\nb = 40\nc = 72\nk = 8\n"
    c = Contraction.from_tccg(text)
    plan = build_vector_plan(c.problem, F32, 0)
    src = codegen.emit_source(plan, text, "gett")
    assert '#include "gett_vector.hpp"' in src
    assert "tccg::gett_vector_kernel<tccg::F32, 1024, 16, 8, 4>" in src
    assert "C must be 16-byte aligned" in src
    plan1 = build_vector_plan(_contraction(3).problem, F64, 0)
    assert plan1.modeA == FREE1
    assert "gett_vector_kernel<tccg::F64, 1024, 16, 8, 1>" in codegen.emit_source(plan1, text, "gett")
