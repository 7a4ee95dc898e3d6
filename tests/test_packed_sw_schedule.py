"""CPU model of the barrier protocol of gett_packed_sw_kernel
(tccg_amd/csrc/gett_packed_sw.hpp), replayed for the compute waves and the
store waves with the kernel's loop structure, over 1..9 tiles of 1..6
k-steps, for both ring depths (double-buffered fragments: copies S steps
ahead; FP8: S - 1).

Intervals are numbered by the barriers passed: interval b lies between
barrier b - 1 and barrier b.  Checked:
- both roles pass the same number of barriers (a mismatch would leave a
  barrier waiting for ever);
- store waves read staged tile i in an interval after the barrier that
  follows its write, and compute waves write staged tile i + 1 only after
  the interval in which tile i was read;
- the C offsets of tile i were copied by compute wave 0 before a copy that
  a barrier ahead of the read waited for (so they have landed), and the
  buffer is not copied into again before the read's interval has ended.
"""
import itertools

import pytest


def compute_role(KT, ntiles, S, db, nbuf):
    ahead = S if db else S - 1
    b = 0                      # barriers passed
    issued = []                # issue call -> (interval, offsets tile or None)
    cursor = {"tile": 0, "k": 0, "end": False}
    waits = []                 # (barrier index, issue call awaited)
    writes = {}

    def issue():
        off = cursor["tile"] if (cursor["k"] == 0 and not cursor["end"]) else None
        issued.append((b, off))
        if cursor["k"] + 1 < KT:
            cursor["k"] += 1
        elif cursor["tile"] + 1 < ntiles:
            cursor["tile"] += 1
            cursor["k"] = 0
        else:
            cursor["end"] = True

    def barrier(awaited=None):
        nonlocal b
        if awaited is not None:
            waits.append((b, awaited))
        b += 1

    for _ in range(ahead):
        issue()
    if db:
        barrier(awaited=0)
    g = 0
    for i in range(ntiles):
        for s in range(KT):
            barrier(awaited=g + 1 if db else g)
            issue()
            g += 1
        if KT == 1:
            barrier()
        writes[i] = b
        barrier()
    return b, writes, issued, waits


def store_role(KT, ntiles, db):
    """reads[i] = (first, last) interval in which staged tile i is read: the
    store waves spread a tile over the intervals after the barriers of k-steps
    0 .. KT-2 of the next tile (one interval when KT == 1)."""
    b = 0
    reads = {}
    after_first = (KT - 1) + (1 if KT == 1 else 0) + 1
    chunks = KT - 1 if KT > 1 else 1
    if db:
        b += 1
    for i in range(ntiles):
        b += 1
        first = b
        for j in range(after_first):
            if i > 0 and j < chunks:
                reads[i - 1] = (first, b)
            b += 1
    if ntiles:
        reads[ntiles - 1] = (b, b)
    return b, reads


@pytest.mark.parametrize("S,db", [(3, True), (4, True), (2, False), (3, False), (4, False)])
def test_store_wave_protocol(S, db):
    nbuf = S + 2
    for KT, ntiles in itertools.product(range(1, 7), range(1, 10)):
        nb_c, writes, issued, waits = compute_role(KT, ntiles, S, db, nbuf)
        nb_s, reads = store_role(KT, ntiles, db)
        assert nb_c == nb_s, (KT, ntiles)
        for i in range(ntiles):
            assert reads[i][0] > writes[i], ("read before publish", KT, ntiles, i)
            if i + 1 < ntiles:
                assert writes[i + 1] > reads[i][1], ("overwrite while read", KT, ntiles, i)
        # offsets of tile i: the issue call that copied them
        dma = {off: (call, iv) for call, (iv, off) in enumerate(issued) if off is not None}
        assert sorted(dma) == list(range(ntiles))
        for i in range(ntiles):
            call, iv = dma[i]
            # a barrier before the read interval waited for this call or a later one
            assert any(bar < reads[i][0] and aw >= call for bar, aw in waits), ("offsets not landed", KT, ntiles, i)
            if i + nbuf in dma:
                assert dma[i + nbuf][1] > reads[i][1], ("offset buffer reused while read", KT, ntiles, i)


def test_model_catches_a_missing_extra_barrier():
    """Without the extra barrier of single-step tiles the next staged tile
    would be written in the interval in which the store waves read the
    previous one."""
    KT, ntiles, S = 1, 3, 3
    b = 1
    writes, reads = {}, {}
    for i in range(ntiles):     # compute: one step barrier, write, publish
        b += 1
        writes[i] = b
        b += 1
    nb_s, reads = store_role(KT, ntiles, True)
    assert any(writes[i + 1] <= reads[i][1] for i in range(ntiles - 1)) or b != nb_s
