"""Host side of the grouped packed GETT (tccg_amd/group.py): operand sharing,
the records the two launches read (checked against the plans they come
from), the work list (every tile of every problem exactly once, longest k
ranges first on each XCD), and the workspace layout (planes disjoint, the
prefetch tail pad after the last one)."""
import ctypes

import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.batch import random_batch
from tccg_amd.group import PACK_TF, PACK_TK, GettGroup, _COperand, _CProblem, group_for_plans
from tccg_amd.plan import F32, F64, PACKED_TAIL_BYTES, PACKED_TILES, build_packed_plan


class FakeTensor:
    def __init__(self, ptr):
        self.ptr = ptr

    def data_ptr(self):
        return self.ptr


def test_record_layouts():
    assert ctypes.sizeof(_COperand) == 48 and ctypes.sizeof(_CProblem) == 72
    assert _CProblem.alpha.offset == 56 and _COperand.fBlocks.offset == 32


@pytest.mark.parametrize("prec,tile", [(F32, 2), (F32, 3), (F64, 0)])
def test_batch_records(prec, tile):
    _check_records(random_batch(40, seed=3), prec, tile)


@pytest.mark.parametrize("batch", range(6))
def test_random_case_records(batch):
    """Batches of eight of the GPU sweep's random contractions
    (tests/random_cases.py: any of m, n, k empty, shuffled index orders,
    ragged extents) on the grouped launch's records."""
    from .random_cases import random_contraction
    _check_records([random_contraction(8 * batch + i) for i in range(8)], F32, 2)


def _check_records(cs, prec, tile):
    plans = [build_packed_plan(c.problem, prec, tile) for c in cs]
    As = [FakeTensor(0x10000000 * (i + 1)) for i in range(len(cs))]
    Bs = [FakeTensor(0x10000000 * (i + 1) + 0x4000000) for i in range(len(cs))]
    Cs = [FakeTensor(0x10000000 * (i + 1) + 0x8000000) for i in range(len(cs))]
    g = group_for_plans(plans, As, Bs, Cs, [1.5] * len(cs), [0.25] * len(cs))
    tables, ops, jobs, probs, work, need = g.records(1 << 40)
    assert need == g.workspace_bytes()
    BM, BN, BK = PACKED_TILES[prec][tile]
    # planes: disjoint, 256-byte aligned, inside the workspace before the tail pad
    spans = sorted((ops[i].out, ops[i].out + g.plane_bytes(g._ops[i])) for i in range(len(g._ops)))
    assert all(a % 256 == 0 for a, _ in spans)
    assert all(b0 <= a1 for (_, b0), (a1, _) in zip(spans, spans[1:]))
    assert spans[-1][1] + PACKED_TAIL_BYTES <= need
    for j, (p, c) in enumerate(zip(plans, cs)):
        q = probs[j]
        assert (q.tilesM, q.tilesN, q.kTiles) == (p.Mp // BM, p.Np // BN, p.Kp // BK)
        assert q.C == Cs[j].ptr and (q.alpha, q.beta) == (1.5, 0.25)
        # the C tables are the plan's
        cm = (q.offCm - (1 << 40)) // 4
        assert np.array_equal(tables[cm:cm + p.Mp].view(np.uint32), p.offCm)
        a_ops = [i for i in range(len(g._ops)) if ops[i].out == q.packedA]
        assert len(a_ops) == 1 and not ops[a_ops[0]].isB
        oa = ops[a_ops[0]]
        assert oa.X == (Bs[j] if p.problem.swapped else As[j]).ptr
        assert oa.fBlocks * PACK_TF == p.Mp and oa.kTiles == p.Kp // BK
    # pack jobs: every 64 x 32 block of every operand once
    want = {(i, f, k) for i in range(len(g._ops)) for f in range(ops[i].fBlocks)
            for k in range(ops[i].kTiles * BK // PACK_TK)}
    got = [tuple(r[:3]) for r in jobs]
    assert len(got) == len(want) and set(got) == want
    for order in ("xcd", "lpt"):
        g.order = order
        work = g.work_list()
        # work: every tile once; padding entries (problem -1) only after a run's end
        want_t = {(j, tm, tn) for j in range(len(cs)) for tm in range(probs[j].tilesM)
                  for tn in range(probs[j].tilesN)}
        got_t = [tuple(r[:3]) for r in work if r[0] >= 0]
        assert len(got_t) == len(want_t) and set(got_t) == want_t
        if order == "xcd":
            assert len(work) % 8 == 0
            # on each XCD (position mod 8) the k ranges never grow along the list
            for x in range(8):
                run = work[x::8, 0]
                valid = run >= 0
                assert np.all(valid[:valid.sum()])
                kt = np.array([probs[int(j)].kTiles for j in run[valid]])
                assert np.all(np.diff(kt) <= 0)
            # the runs carry about equal work
            per = [sum(probs[int(j)].kTiles for j in work[x::8, 0] if j >= 0) for x in range(8)]
            assert max(per) - min(per) <= max(probs[j].kTiles for j in range(len(cs))) * 2
        else:
            kt = np.array([probs[int(j)].kTiles for j in work[:, 0]])
            assert np.all(np.diff(kt) <= 0) and np.all(work[:, 0] >= 0)


def test_walk_and_xcd_chunks():
    """Inside a problem the tiles follow tile_coords' walk (groups of 8 tile
    rows, column by column), and in "lpt" order the positions of one XCD
    hold a contiguous range of that walk."""
    from tccg_amd.group import _walk
    w = _walk(19, 5)
    assert len(set(w)) == 95 and w[:9] == [(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (7, 0), (0, 1)]
    assert w[80:83] == [(16, 0), (17, 0), (18, 0)]
    g = GettGroup(F32, PACKED_TILES[F32].index((128, 128, 16)), order="lpt")
    offK = np.arange(64, dtype=np.uint32)
    a = g.operand(1, np.arange(19 * 128, dtype=np.uint32), offK, 0, False)
    b = g.operand(2, np.arange(5 * 128, dtype=np.uint32), offK, 0, True)
    g.problem(a, b, 3, np.arange(19 * 128, dtype=np.uint32), np.arange(5 * 128, dtype=np.uint32), 1.0, 0.0)
    work = g.work_list()
    index = {t: i for i, t in enumerate(w)}
    for x in range(8):
        logical = [index[(int(r[1]), int(r[2]))] for r in work[x::8]]
        assert logical == list(range(logical[0], logical[0] + len(logical)))


def test_operands_shared_by_key():
    g = GettGroup(F32, PACKED_TILES[F32].index((128, 128, 16)))
    offF = np.arange(128, dtype=np.uint32)
    offK = np.arange(64, dtype=np.uint32)
    a0 = g.operand(1000, offF, offK, 0, False, key=("A", 0))
    a1 = g.operand(1000, offF, offK, 0, False, key=("A", 0))
    b0 = g.operand(2000, offF, offK, 0, True, key=("B", 0))
    assert a0 == a1 and a0 != b0 and len(g._ops) == 2
    for c in range(5):
        g.problem(a0, b0, 3000 + 4 * c, offF, offF, 1.0, 0.0)
    _, ops, jobs, probs, work, need = g.records(0)
    assert len(jobs) == 2 * (128 // PACK_TF) * (64 // PACK_TK)      # two operands packed once
    assert all(probs[j].packedA == ops[a0].out for j in range(5))
    with pytest.raises(ValueError):
        g.operand(1000, np.arange(100, dtype=np.uint32), offK, 0, False)     # rows not a multiple of 64 / BM
    with pytest.raises(ValueError):
        g.problem(b0, a0, 0, offF, offF, 1.0, 0.0)                          # operand roles swapped
