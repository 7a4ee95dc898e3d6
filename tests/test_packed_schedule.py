"""CPU model of the copy/wait schedule of gett_packed_kernel
(tccg_amd/csrc/gett_packed.hpp): for every k-step the copies whose LDS stage
the step reads must have landed when its barrier passes.

A wave's vector-memory operations complete in issue order (vmcnt), so
s_waitcnt vmcnt(N) guarantees every operation except the last N issued.  The
model replays one wave's issue order -- ring copies (PER_WAVE per k-step,
S stages, running S steps ahead with fragment double-buffering and S - 1
without), the epilogue's C-offset loads, the drain wait of a persistent
epilogue and the C stores -- over persistent walks of several tiles with
1..7 k-steps, with any number of offset loads and of stores (a wave whose
sub-tile is all padding stores nothing), and checks each step's wait.
"""
import itertools

import pytest


def replay(S, per_wave, db, kts, offset_loads, stores, persist=True, drain=True, fresh_wait=63):
    """Returns the list of (global step, ok) for every k-step of the walk."""
    ahead = S if db else S - 1
    ops = []          # issue order: ("copy", issue call) / ("load",) / ("store",)
    done = 0          # ops[:done] have completed (a drain or a wait)
    issued = [0]

    def issue():
        # call i fills the stage that global step i reads (past the last step
        # the cursor stays put: a re-read into a stage nobody reads)
        ops.extend([("copy", issued[0])] * per_wave)
        issued[0] += 1

    def wait(n):
        nonlocal done
        done = max(done, len(ops) - n)

    def landed(step):
        idx = [i for i, op in enumerate(ops) if op == ("copy", step)]
        return bool(idx) and max(idx) < done

    results = []
    for _ in range(ahead):
        issue()
    if db:
        wait((S - 1) * per_wave)
        results.append((-1, landed(0)))
    fresh = 0
    g = 0
    for t, kt in enumerate(kts):
        for s in range(kt):
            awaited = g + 1 if db else g
            if persist and fresh > 0:
                wait(fresh_wait)
                fresh -= 1
            else:
                wait((S - 2) * per_wave)
            results.append((g, landed(awaited)))
            issue()
            g += 1
        ops.extend([("load",)] * offset_loads)
        if drain:
            wait(0)
        ops.extend([("store",)] * stores)
        fresh = S - 1
    return results


# (S, copies per wave and k-step) of the packed tiles: FP32 128x128x16 / 192x192x16,
# split-BF16 192x192x16, FP8 192x192x128 (no double-buffering)
SHAPES = [(3, 4, True), (3, 6, True), (3, 9, True), (3, 12, False), (4, 4, True)]


@pytest.mark.parametrize("S,per_wave,db", SHAPES)
def test_every_step_waits_for_its_copies(S, per_wave, db):
    for kts in itertools.chain(itertools.product(range(1, 8), repeat=2), [(1, 1, 1, 1), (2, 1, 5, 1, 3)]):
        for loads in (0, 3, 15, 51):
            for stores in (0, 1, 48, 300):
                res = replay(S, per_wave, db, list(kts), loads, stores)
                bad = [g for g, ok in res if not ok]
                assert not bad, (S, per_wave, db, kts, loads, stores, bad)


def test_model_catches_a_missing_drain():
    """Without the drain, the no-wait steps after an epilogue would read copies
    that may still be in flight behind the offset loads: the model sees it."""
    res = replay(3, 4, True, [4, 4], offset_loads=3, stores=0, drain=False)
    assert not all(ok for _, ok in res)


def test_model_catches_an_early_fresh_count():
    """The count the kernel used before the drain (copies + 51 offset loads)
    is wrong when the compiler issues fewer, wider offset loads."""
    res = replay(3, 6, True, [3, 3, 3], offset_loads=15, stores=0, drain=False,
                 fresh_wait=(3 - 2) * 6 + 3 + 3 * 16)
    assert not all(ok for _, ok in res)
