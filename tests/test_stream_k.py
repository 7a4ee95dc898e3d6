"""The stream-k walk's partition and partial-sum protocol, modelled on the CPU.

gett_packed_sk_kernel (tccg_amd/csrc/gett_packed.hpp) walks all but the last
one or two rounds of tiles whole (tile g + i * grid of the XCD-chunked order
for workgroup g), splits the k-steps of the remaining tiles of each XCD into
equal segments, one per workgroup, walks each segment's pieces from its last
tile back to its first, publishes the raw sums of a piece that ends inside
its tile and lets the piece holding a tile's last k-step add the others.  The model below restates that arithmetic (the same
integer divisions) and checks what the kernel relies on: every (tile, k-step)
unit is computed exactly once, every cut tile is finished by exactly one
combining piece that adds every other piece of it, every wait is for a lower
workgroup index on the same XCD, and a workgroup publishes (at most once)
before it waits for anything, so the waits end.  No GPU is used.
"""
import ctypes
import itertools

import pytest

from tccg_amd import build
from tccg_amd.plan import (F32, F64, BF16X3, MI355X, PACKED_TILES, build_packed_plan, candidate_plans,
                           plan_from_kind, plan_kind)

NX = 8
TAIL_ROUNDS = 1   # kSkTailRounds


def xcd_chunked(bid, nblocks):
    """The kernels' map of a block id to its XCD's contiguous chunk."""
    q, r = nblocks // NX, nblocks % NX
    xcd, local = bid % NX, bid // NX
    return xcd * (q + 1) + local if xcd < r else r * (q + 1) + (xcd - r) * q + local


def dp_tiles(grid, ntiles):
    rounds = ntiles // grid
    return (rounds - TAIL_ROUNDS) * grid if rounds > TAIL_ROUNDS else 0


def segments(grid, ntiles, KT):
    """{workgroup: (xcd, seg, nseg, units, T0, s, e)} of the stream-k tiles,
    as the kernel computes them."""
    out = {}
    nx = min(NX, grid)   # XCD groups: one per workgroup of a grid under 8
    dp = dp_tiles(grid, ntiles)
    tail = ntiles - dp
    for g in range(grid):
        x, seg = g % nx, g // nx
        nseg = (grid - x + nx - 1) // nx
        T0, T1 = dp + x * tail // nx, dp + (x + 1) * tail // nx
        units = (T1 - T0) * KT
        s, e = seg * units // nseg, (seg + 1) * units // nseg
        out[g] = (x, seg, nseg, units, T0, s, e)
    return out


def pieces(grid, ntiles, KT):
    """Per workgroup, its pieces in walk order: (tile, k0, nk, produce, combine, gathered slots)."""
    segs = segments(grid, ntiles, KT)
    dp = dp_tiles(grid, ntiles)
    walk = {}
    for g, (x, seg, nseg, units, T0, s, e) in segs.items():
        lst = [(xcd_chunked(g + i * grid, dp), 0, KT, False, False, []) for i in range(dp // grid)]
        if s < e:
            t = (e - 1) // KT
            while t >= s // KT:
                lo, hi = max(s, t * KT), min(e, (t + 1) * KT)
                produce = hi < (t + 1) * KT
                combine = lo > t * KT and not produce
                gathered = []
                if combine:
                    first = t * KT
                    for jj in range(seg - 1, -1, -1):
                        ej = (jj + 1) * units // nseg
                        if ej <= first:
                            break
                        if jj * units // nseg >= ej:
                            continue
                        gathered.append(x + min(NX, grid) * jj)
                lst.append((T0 + t, lo - t * KT, hi - lo, produce, combine, gathered))
                t -= 1
        walk[g] = lst
    return walk


CASES = [(256, 729, 144), (512, 729, 144), (256, 256, 8), (256, 300, 3), (8, 1, 5), (16, 3, 1), (1, 3, 4), (3, 2, 5),
         (256, 20, 2), (24, 17, 7), (40, 100, 9), (7, 5, 4), (512, 324, 324), (256, 1000, 11), (8, 100, 3),
         (37, 200, 5), (256, 729, 72)]


@pytest.mark.parametrize("grid,ntiles,KT", CASES)
def test_units_covered_once_and_protocol(grid, ntiles, KT):
    walk = pieces(grid, ntiles, KT)
    seen = {}
    holders = {}
    for g, lst in walk.items():
        for (tile, k0, nk, produce, combine, gathered) in lst:
            assert nk >= 1 and 0 <= k0 and k0 + nk <= KT
            for k in range(k0, k0 + nk):
                assert (tile, k) not in seen, "unit computed twice"
                seen[(tile, k)] = g
            holders.setdefault(tile, []).append((g, k0, nk, produce, combine, gathered))
    assert len(seen) == ntiles * KT
    published = {}
    ndp = dp_tiles(grid, ntiles) // grid
    for g, lst in walk.items():
        prod = [i for i, pc in enumerate(lst) if pc[3]]
        # at most one published partial per workgroup (its slot), and it is the
        # first stream-k piece the workgroup walks (after the same number of
        # whole tiles in every workgroup), before any wait
        assert len(prod) <= 1
        assert all(not pc[3] and not pc[4] and pc[1:3] == (0, KT) for pc in lst[:ndp])
        if prod:
            assert prod[0] == ndp
            published[g] = lst[ndp][0]
        for i, pc in enumerate(lst):
            if pc[4]:
                assert i == len(lst) - 1, "a combining piece is the workgroup's last"
    for tile, hs in holders.items():
        if len(hs) == 1:
            (g, k0, nk, produce, combine, gathered) = hs[0]
            assert (k0, nk) == (0, KT) and not produce and not combine
            continue
        finishers = [h for h in hs if h[1] + h[2] == KT]
        assert len(finishers) == 1
        g, k0, nk, produce, combine, gathered = finishers[0]
        assert combine and not produce and k0 > 0
        others = sorted(h[0] for h in hs if h[0] != g)
        assert sorted(gathered) == others, "the finisher adds exactly the tile's other pieces"
        for src in gathered:
            nx = min(NX, grid)
            assert src < g and src % nx == g % nx, "waits only for lower workgroups on its XCD"
            assert published.get(src) == tile, "each waited-for workgroup publishes this tile"
        for h in hs:
            if h[0] != g:
                assert h[3] and not h[4]


def test_each_partial_has_exactly_one_reader():
    for grid, ntiles, KT in CASES:
        walk = pieces(grid, ntiles, KT)
        readers = {}
        for g, lst in walk.items():
            for pc in lst:
                for src in pc[5]:
                    readers[src] = readers.get(src, 0) + 1
        producers = [g for g, lst in walk.items() if any(pc[3] for pc in lst)]
        assert sorted(readers) == sorted(producers)
        assert all(v == 1 for v in readers.values()), "a flag is cleared by its one reader"


def test_data_parallel_rounds():
    """All but the last one or two rounds of tiles are walked whole, the same
    number by every workgroup; the chunked order covers them once."""
    for grid, ntiles, KT in CASES:
        dp = dp_tiles(grid, ntiles)
        assert dp % grid == 0 and (dp == 0 or grid <= ntiles - dp < 2 * grid)
        assert sorted(xcd_chunked(v, dp) for v in range(dp)) == list(range(dp))


def test_balanced_segments():
    """No workgroup holds more than one k-step above the even share."""
    for grid, ntiles, KT in CASES:
        segs = segments(grid, ntiles, KT)
        for x in range(min(NX, grid)):
            lens = [e - s for (xx, _, _, _, _, s, e) in segs.values() if xx == x]
            if lens:
                assert max(lens) - min(lens) <= 1


def test_stream_k_plans_and_kinds():
    from tccg_amd import Contraction
    from tccg_amd.workloads import get
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    for prec in (F32, F64, BF16X3):
        p = build_packed_plan(c.problem, prec, 0, stream_k=2)
        assert p.walk == 5 and p.grid == 2 * MI355X["cus"] and "streamk grid=512" in p.describe()
        q = plan_from_kind(c.problem, prec, 0, False, plan_kind(p))
        assert (q.walk, q.grid, q.ksplit) == (5, 512, 1)
    with pytest.raises(ValueError):
        build_packed_plan(c.problem, F32, 0, persistent=True, stream_k=1)
    with pytest.raises(ValueError):
        build_packed_plan(c.problem, F32, 0, stream_k=5)
    # offered where whole rounds of tiles leave the last one partly idle
    walks = [(p.walk, p.grid) for p in candidate_plans(c.problem, F32)]
    assert (5, 256) in walks and (5, 512) in walks
    # not for a contraction whose tiles and k-steps give the grid little work
    small = Contraction.from_tccg(get("cpu48"))
    assert all(p.walk != 5 for p in candidate_plans(small.problem, F32))
    # offered under one round of tiles when k is long (next to split-k): the
    # first adjoint of abcd-aklb-lckd has 162 tiles of 192 x 192, k = 5184
    from tccg_amd.autograd import adjoints
    dA, _ = adjoints(Contraction.from_tccg(get("abcd-aklb-lckd")))
    plans = candidate_plans(dA.problem, F32)
    assert any(p.walk == 5 and p.ntiles() < MI355X["cus"] for p in plans)
    assert any(p.ksplit > 1 for p in plans)


def test_stream_k_workspace_bytes_agree():
    """Partials after the packed operands, then the flags: the library's size
    and the plan's agree (host arithmetic only)."""
    path = build.library_path()
    if not path.exists():
        pytest.skip("library not built")
    from tccg_amd import Contraction
    from tccg_amd.runtime import _CPlan, load_library
    from tccg_amd.workloads import get
    lib = load_library()
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    for prec, tiles in PACKED_TILES.items():
        for t, spc in itertools.product(range(len(tiles)), (1, 2)):
            p = build_packed_plan(c.problem, prec, t, stream_k=spc)
            cp = _CPlan()
            cp.Mp, cp.Np, cp.Kp, cp.tile, cp.walk, cp.grid = p.Mp, p.Np, p.Kp, p.tile, p.walk, p.grid
            n = lib.tccg_packed_workspace_bytes(prec, ctypes.byref(cp))
            esz = 8 if prec == F64 else 4
            assert n == p.workspace_bytes() >= p.grid * p.BM * p.BN * esz + 4 * (p.grid + 1)


@pytest.mark.parametrize("seed", range(48))
def test_random_case_partitions(seed):
    """The (grid, tiles, k-steps) triples of the GPU sweep's stream-k plans
    (tests/random_cases.py, packed tile 0, grids 3, 37 and 256) through the
    partition and protocol checks above."""
    from tccg_amd.plan import BF16X3
    from .random_cases import random_contraction
    c = random_contraction(seed)
    for prec in (F32, F64, BF16X3):
        p = build_packed_plan(c.problem, prec, 0, stream_k=1)
        for grid in (3, 37, 256):
            test_units_covered_once_and_protocol(grid, p.ntiles(), p.Kp // p.BK)
