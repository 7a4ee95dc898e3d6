"""Adjoint derivation and batch bookkeeping (CPU; no kernel launches)."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref
from tccg_amd import Contraction
from tccg_amd.autograd import adjoints
from tccg_amd.batch import BATCH_SHAPES, GettBatch, pattern_key, random_batch
from tccg_amd.plan import F32, TILES
from tccg_amd.workloads import BATCH_PATTERNS, WORKLOADS


def labels(T):
    return [i.label for i in T.indices]


def test_adjoint_index_lists():
    c = Contraction.from_tccg(WORKLOADS["abcd-aklb-lckd"])
    dA, dB = adjoints(c)
    assert labels(dA.C0) == ["a", "k", "l", "b"] and labels(dA.A0) == ["a", "b", "c", "d"] and labels(dA.B0) == ["l", "c", "k", "d"]
    assert labels(dB.C0) == ["l", "c", "k", "d"] and labels(dB.A0) == ["a", "k", "l", "b"] and labels(dB.B0) == ["a", "b", "c", "d"]
    # dA contracts over the forward's n set, dB over its m set
    assert sorted(i.label for i in dA.problem.k) == ["c", "d"]
    assert sorted(i.label for i in dB.problem.k) == ["a", "b"]


def test_adjoints_are_gradients_numpy():
    """<dC, d/dA (A.B)> identity: sum(dC * C(A,B)) == sum(dA * A) for bilinear C."""
    rng = np.random.default_rng(0)
    sizes = dict(a=3, b=4, c=2, d=5, k=3, l=2)
    c = Contraction.from_labels(["a", "b", "c", "d"], ["a", "k", "l", "b"], ["l", "c", "k", "d"], sizes, alpha=1.7)
    A = rng.standard_normal(c.A0.torch_shape())
    B = rng.standard_normal(c.B0.torch_shape())
    dC = rng.standard_normal(c.C0.torch_shape())
    dAc, dBc = adjoints(c)
    C = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, alpha=1.7)
    dA = contract_ref(labels(dAc.C0), labels(dAc.A0), labels(dAc.B0), dC, B, alpha=1.7)
    dB = contract_ref(labels(dBc.C0), labels(dBc.A0), labels(dBc.B0), A, dC, alpha=1.7)
    assert np.isclose(np.sum(dC * C), np.sum(dA * A))
    assert np.isclose(np.sum(dC * C), np.sum(dB * B))


def test_random_batch_extents():
    cs = random_batch(64, seed=3)
    for c in cs:
        for T in (c.A0, c.B0, c.C0):
            for i in T.indices:
                assert i.size % 8 == 0 and 8 <= i.size <= 96
    keys = {pattern_key(c) for c in cs}
    assert len(keys) == len(BATCH_PATTERNS)


def test_batch_shapes_match_tiles():
    assert TILES[F32][BATCH_SHAPES[0]] == (64, 64, 32) and TILES[F32][BATCH_SHAPES[1]] == (128, 128, 32)
    b = GettBatch(random_batch(8, seed=1))
    for p in b.plans:
        assert p.tile in BATCH_SHAPES.values()
