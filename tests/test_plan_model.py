"""CPU model of the table-driven addressing every GETT kernel shares, over
random contractions and every candidate plan (gather, blocked, persistent,
streaming, packed, vector): element (m, k) of the internal A operand is read
at offAm[m] + offAk[k], B's (n, k) at offBn[n] + offBk[k], and C's (m, n)
written at offCm[m] + offCn[n]; positions marked INVALID are read as zero and
never written.  The product computed through the tables must equal the FP64
einsum oracle, every valid access must be in bounds, the validity of A's and
C's m tables (B's and C's n tables, A's and B's k tables) must agree, and every
element of C must be written exactly once."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.plan import F32, F64, INVALID, candidate_plans

INV = np.int64(INVALID)


def random_contraction(rng):
    labels = iter("abcdefghij")
    m = [next(labels) for _ in range(rng.integers(1, 3))]
    n = [next(labels) for _ in range(rng.integers(1, 3))]
    k = [next(labels) for _ in range(rng.integers(1, 3))]
    sizes = {l: int(rng.integers(1, 10)) for l in m + n + k}
    perm = lambda xs: [xs[i] for i in rng.permutation(len(xs))]
    return Contraction.from_labels(perm(m + n), perm(m + k), perm(n + k), sizes)


def emulate(plan, A, B, C, alpha, beta):
    pb = plan.problem
    opA, opB = (B, A) if pb.swapped else (A, B)
    a, b, c = opA.reshape(-1), opB.reshape(-1), C.reshape(-1).copy()
    t = {key: getattr(plan, key).astype(np.int64) for key in ("offAm", "offAk", "offBn", "offBk", "offCm", "offCn")}
    vm, vn, vk = t["offAm"] != INV, t["offBn"] != INV, t["offAk"] != INV
    assert np.array_equal(vm, t["offCm"] != INV) and np.array_equal(vn, t["offCn"] != INV)
    assert np.array_equal(vk, t["offBk"] != INV)
    am, ak, bn, bk = t["offAm"][vm], t["offAk"][vk], t["offBn"][vn], t["offBk"][vk]
    cm, cn = t["offCm"][vm], t["offCn"][vn]
    # bounds of every valid access
    assert am.max() + ak.max() < a.size and bn.max() + bk.max() < b.size
    assert cm.max() + cn.max() < c.size
    # every element of C written exactly once
    idx = (cm[:, None] + cn[None, :]).reshape(-1)
    assert idx.size == c.size and np.unique(idx).size == c.size
    Am = a[am[:, None] + ak[None, :]]
    Bm = b[bn[:, None] + bk[None, :]]
    out = alpha * (Am @ Bm.T)
    cidx = cm[:, None] + cn[None, :]
    c[cidx] = out + beta * c[cidx]
    return c.reshape(C.shape)


@pytest.mark.parametrize("seed", range(24))
def test_every_plan_addresses_the_contraction(seed):
    rng = np.random.default_rng(seed)
    c = random_contraction(rng)
    sa, sb, sc = c.shapes()
    A, B, C = rng.uniform(-1, 1, sa), rng.uniform(-1, 1, sb), rng.uniform(-1, 1, sc)
    labels = lambda T: [i.label for i in T.indices]
    ref = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, C, 1.25, -0.5)
    prec = F64 if seed % 2 else F32
    plans = candidate_plans(c.problem, prec)
    assert plans
    kinds = set()
    for plan in plans:
        kinds.add((plan.packed, bool(plan.grid), plan.walk))
        out = emulate(plan, A, B, C, 1.25, -0.5)
        assert rel_frobenius(out, ref) < 1e-13, (str(c), plan.describe())
    assert any(p for p, _, _ in kinds) and any(not p for p, _, _ in kinds), kinds


@pytest.mark.parametrize("seed", range(48))
def test_random_index_sets_every_plan(seed):
    """The random cases of the GPU sweep (any of the m, n, k sets may be
    empty, index orders shuffled, ragged extents): every candidate plan's
    tables address the contraction, and the internal A operand holds C's
    stride-one index (with an empty n set it may be either input)."""
    from .random_cases import random_contraction
    c = random_contraction(seed)
    rng = np.random.default_rng(seed)
    sa, sb, sc = c.shapes()
    A, B, C = rng.uniform(-1, 1, sa), rng.uniform(-1, 1, sb), rng.uniform(-1, 1, sc)
    labels = lambda T: [i.label for i in T.indices]
    ref = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, C, 1.25, -0.5)
    pb = c.problem
    assert pb.opA.has_index(pb.C.indices[0]) and not pb.opB.has_index(pb.C.indices[0])
    for prec in (F32, F64):
        for plan in candidate_plans(pb, prec):
            out = emulate(plan, A, B, C, 1.25, -0.5)
            assert rel_frobenius(out, ref) < 1e-12, (str(c), plan.describe())


def test_vector_operand_in_either_input():
    """C[j] = A[k] B[k, j] and C[j] = A[k, j] B[k]: the matrix operand is the
    internal A operand in both, whichever input it is."""
    for c_l, a_l, b_l in ((["j"], ["k"], ["k", "j"]), (["j"], ["k", "j"], ["k"]),
                          (["j", "i"], ["k", "h"], ["h", "i", "k", "j"])):
        c = Contraction.from_labels(c_l, a_l, b_l, {"j": 37, "k": 21, "i": 5, "h": 3})
        pb = c.problem
        assert pb.opA.has_index(pb.C.indices[0]) and pb.N == 1, str(pb)
        assert pb.swapped == ("j" not in a_l)
        rng = np.random.default_rng(len(a_l))
        sa, sb, sc = c.shapes()
        A, B, C = rng.uniform(-1, 1, sa), rng.uniform(-1, 1, sb), rng.uniform(-1, 1, sc)
        labels = lambda T: [i.label for i in T.indices]
        ref = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, C, 2.0, 0.5)
        for plan in candidate_plans(pb, F32):
            assert rel_frobenius(emulate(plan, A, B, C, 2.0, 0.5), ref) < 1e-12, plan.describe()


def emulate_store_waves(plan, C_exact, C_init):
    """The store-wave kernel's write of a finished tile: in run order (crun
    R > 0) the group of four m positions m4 of a run starting at r0 goes to
    offCm[r0] + (m4 - r0) + offCn[n]; with cvec, aligned groups of four m
    positions are written as four consecutive elements from offCm[m4] +
    offCn[n] (16-byte aligned); otherwise element by element."""
    cm = plan.offCm.astype(np.int64)
    cn = plan.offCn.astype(np.int64)
    out = C_init.reshape(-1).copy()
    exact = C_exact.reshape(-1)
    R = plan.crun[0]
    if R:
        assert plan.BM % R == 0 and R % 4 == 0
        for n in range(cn.size):
            if cn[n] == INV:
                continue
            for r0 in range(0, cm.size, R):
                if cm[r0] == INV:
                    continue
                for m4 in range(r0, r0 + R, 4):
                    base = cm[r0] + (m4 - r0) + cn[n]
                    assert base % 4 == 0
                    for j in range(4):
                        out[base + j] = exact[cm[m4 + j] + cn[n]]
        return out.reshape(C_init.shape)
    for n in range(cn.size):
        if cn[n] == INV:
            continue
        for m4 in range(0, cm.size, 4):
            if plan.cvec:
                if cm[m4] == INV:
                    continue
                base = cm[m4] + cn[n]
                assert base % 4 == 0
                # the value of position (m4 + j, n) lands at base + j
                for j in range(4):
                    out[base + j] = exact[cm[m4 + j] + cn[n]]
            else:
                for j in range(4):
                    if cm[m4 + j] != INV:
                        out[cm[m4 + j] + cn[n]] = exact[cm[m4 + j] + cn[n]]
    return out.reshape(C_init.shape)


@pytest.mark.parametrize("seed", range(12))
def test_store_wave_groups(seed):
    """Every store-wave plan's C writes (16-byte groups when cvec) land on the
    right elements; contractions whose C has a stride-one extent divisible by
    4 get cvec plans."""
    from tccg_amd.plan import BF16X3, PACKED_TILES, build_packed_plan, sw_built
    rng = np.random.default_rng(100 + seed)
    c = random_contraction(rng)
    if seed % 3 == 0:   # make C's stride-one extent a multiple of 4
        sizes = {i.label: i.size for T in (c.C0, c.A0, c.B0) for i in T.indices}
        sizes[c.C0.indices[0].label] = 4 * int(rng.integers(1, 4))
        labels = lambda T: [i.label for i in T.indices]
        c = Contraction.from_labels(labels(c.C0), labels(c.A0), labels(c.B0), sizes)
    sc = c.shapes()[2]
    exact = rng.uniform(-1, 1, sc)
    init = np.full(sc, np.nan)
    for prec in (F32, BF16X3):
        for tile in range(len(PACKED_TILES[prec])):
            if not sw_built(prec, tile):
                continue
            plan = build_packed_plan(c.problem, prec, tile, store_waves=True)
            assert plan.walk == 4 and plan.grid > 0
            if seed % 3 == 0:
                assert plan.cvec, plan.describe()
            out = emulate_store_waves(plan, exact, init)
            assert np.array_equal(out, exact), (str(c), plan.describe())


def test_store_wave_run_order():
    """C[a,b,...] with a in the m set and b in the n set: the run order is
    chosen (runs of a's extent) when a's extent divides BM, and its writes
    land on the right elements; the precision contraction's shape gets runs
    of 24 with the 192-row tiles."""
    from tccg_amd.plan import FP8, PACKED_TILES, build_packed_plan, sw_built
    rng = np.random.default_rng(7)
    c = Contraction.from_labels(["a", "b", "c", "d"], ["d", "k", "a"], ["k", "c", "b"],
                                dict(a=8, b=5, c=3, d=7, k=4))
    sc = c.shapes()[2]
    exact = rng.uniform(-1, 1, sc)
    seen = 0
    for prec in (F32, FP8):
        for tile in range(len(PACKED_TILES[prec])):
            if not sw_built(prec, tile):
                continue
            plan = build_packed_plan(c.problem, prec, tile, store_waves=True)
            assert plan.crun[0] == (8 if plan.BM % 8 == 0 else 0), plan.describe()
            assert np.array_equal(emulate_store_waves(plan, exact, np.full(sc, np.nan)), exact)
            seen += 1
    assert seen
    from tccg_amd.workloads import WORKLOADS
    c3 = Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"])
    t = PACKED_TILES[FP8].index((192, 64, 128))
    assert build_packed_plan(c3.problem, FP8, t, store_waves=True).crun[0] == 24


def emulate_split(plan, A, B, C, alpha, beta):
    """Split-k plans: work item (ks, tile) sums k-steps [ks*KT/S, (ks+1)*KT/S)
    of the tile into a partial tile; the reduction adds the parts in order and
    writes alpha * sum + beta * C through C's tables."""
    pb = plan.problem
    opA, opB = (B, A) if pb.swapped else (A, B)
    a, b, c = opA.reshape(-1), opB.reshape(-1), C.reshape(-1).copy()
    t = {key: getattr(plan, key).astype(np.int64) for key in ("offAm", "offAk", "offBn", "offBk", "offCm", "offCn")}
    KT, S = plan.Kp // plan.BK, plan.ksplit
    tilesM, tilesN = plan.Mp // plan.BM, plan.Np // plan.BN
    part = np.zeros((S, tilesM * tilesN, plan.BN, plan.BM))
    covered = np.zeros(KT, dtype=int)
    for ks in range(S):
        k0, k1 = ks * KT // S, (ks + 1) * KT // S
        if ks == 0:
            assert k0 == 0
        covered[k0:k1] += 1
        ks_idx = np.arange(k0 * plan.BK, k1 * plan.BK)
        ak, bk = t["offAk"][ks_idx], t["offBk"][ks_idx]
        vk = ak != INV
        for tm in range(tilesM):
            am = t["offAm"][tm * plan.BM:(tm + 1) * plan.BM]
            for tn in range(tilesN):
                bn = t["offBn"][tn * plan.BN:(tn + 1) * plan.BN]
                Am = np.where((am[:, None] != INV) & vk[None, :], a[np.where(am[:, None] != INV, am[:, None], 0)
                                                                   + np.where(vk, ak, 0)[None, :]], 0.0)
                Bm = np.where((bn[:, None] != INV) & vk[None, :], b[np.where(bn[:, None] != INV, bn[:, None], 0)
                                                                   + np.where(vk, bk, 0)[None, :]], 0.0)
                part[ks, tm * tilesN + tn] = Bm @ Am.T          # rows n of BM
    assert np.all(covered == 1)          # every k-step in exactly one part
    total = part.sum(axis=0)
    for tm in range(tilesM):
        for tn in range(tilesN):
            cm = t["offCm"][tm * plan.BM:(tm + 1) * plan.BM]
            cn = t["offCn"][tn * plan.BN:(tn + 1) * plan.BN]
            for n in range(plan.BN):
                for m in range(plan.BM):
                    if cm[m] != INV and cn[n] != INV:
                        i = cm[m] + cn[n]
                        c[i] = alpha * total[tm * tilesN + tn, n, m] + beta * c[i]
    return c.reshape(C.shape)


@pytest.mark.parametrize("seed", range(6))
def test_split_k_plans(seed):
    from tccg_amd.plan import BF16X3, PACKED_TILES, build_packed_plan
    rng = np.random.default_rng(200 + seed)
    sizes = dict(a=int(rng.integers(3, 40)), b=int(rng.integers(3, 40)), k=int(rng.integers(70, 200)))
    c = Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], sizes)
    sa, sb, sc = c.shapes()
    A, B, C = rng.uniform(-1, 1, sa), rng.uniform(-1, 1, sb), rng.uniform(-1, 1, sc)
    ref = contract_ref(["a", "b"], ["a", "k"], ["k", "b"], A, B, C, 0.75, 1.5)
    for prec in (F32, BF16X3):
        tile = PACKED_TILES[prec].index((128, 128, 16))
        for ks in (2, 3, 5):
            plan = build_packed_plan(c.problem, prec, tile, ksplit=ks)
            assert plan.ksplit == ks and plan.grid == 0
            out = emulate_split(plan, A, B, C, 0.75, 1.5)
            assert rel_frobenius(out, ref) < 1e-13, (sizes, ks)
