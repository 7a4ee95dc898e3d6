"""The restated AVX2 planner's index algebra (oracle/avx2_model.py): prime
factors, block-size candidates, index-set splitting (CPU)."""
import itertools

import numpy as np
import pytest

from oracle.avx2_model import get_possible_values, get_perm, prime_factors, split_index, split_index_set, transpose_args
from tccg_amd.ir import Index, Tensor


@pytest.mark.parametrize("n,f", [(1, [1]), (2304, [2] * 8 + [3] * 2), (5184, [2] * 6 + [3] * 4),
                                 (13824, [2] * 9 + [3] * 3), (192, [2] * 6 + [3]), (97, [97]),
                                 (2 * 5003, [2])])
def test_prime_factors_known(n, f):
    # 5003 > 5000: the reference's trial division stops at 5000 and drops it
    assert prime_factors(n) == f


def test_prime_factors_product():
    for n in range(2, 3000):
        assert int(np.prod(prime_factors(n))) == n


def brute_possible_values(block, must, target, maxn):
    out = []
    a = target // block
    cands = [x * block for x in range(1, a + 1) if must % (x * block) == 0]
    return sorted(cands, reverse=True)[:maxn]


@pytest.mark.parametrize("block,must,target,maxn", [
    (24, 2304, 192, 3), (24, 5184, 192, 3), (4, 2304, 192, 3), (4, 5184, 192, 3), (1, 64, 256, 4),
    (1, 2304, 256, 4), (1, 8, 256, 4), (1, 96, 256, 4), (1, 1024, 256, 4), (24, 147456, 4600, 3),
    (12, 5184, 96, 3)])
def test_possible_values(block, must, target, maxn):
    assert get_possible_values(block, must, target, maxn) == brute_possible_values(block, must, target, maxn)


def test_possible_values_known_answers():
    # hand-derived candidates of the CPU configuration (mc/nc/kc tables)
    assert get_possible_values(24, 2304, 192, 3) == [192, 144, 96]
    assert get_possible_values(1, 64, 256, 4) == [64, 32, 16, 8]
    assert get_possible_values(24, 147456, 4600, 3) == [3072, 2304, 1536]
    assert get_possible_values(1, 96, 256, 4) == [96, 48, 32, 24]
    assert get_possible_values(1, 1024, 256, 4) == [256, 128, 64, 32]


def T(label, spec):
    return Tensor(label, [Index(l, s) for l, s in spec])


def test_split_index():
    i0, i1 = split_index(Index("a", 72), 8)
    assert (i0.label, i0.size, i1.label, i1.size) == ("a_0", 8, "a_1", 9)


def brute_splits(indices, size):
    """All assignments of a divisor d_i of each extent with prod d_i == size."""
    out = set()
    divs = [[d for d in range(1, i.size + 1) if i.size % d == 0] for i in indices]
    for combo in itertools.product(*divs):
        if int(np.prod(combo)) == size:
            out.add(tuple((i.label, d) for i, d in zip(indices, combo) if d > 1))
    return out


@pytest.mark.parametrize("sizes,size", [((72, 72), 192), ((24, 24, 24), 96), ((48, 48), 144),
                                        ((8, 12, 6), 24), ((5, 7), 35), ((16,), 8)])
def test_split_index_set_matches_brute_force(sizes, size):
    labels = "abcdefg"[: len(sizes)]
    idx = [Index(l, s) for l, s in zip(labels, sizes)]
    A = T("A", [(l, s) for l, s in zip(labels, sizes)] + [("k", 3)])
    C = T("C", [(l, s) for l, s in zip(labels, sizes)])
    sols = split_index_set(idx, size, A, False, C, False)
    got = {tuple((i.label.split("_")[0], i.size) for i in ind0) for ind0, _, _, _ in sols}
    got = {tuple(sorted(x)) for x in got}
    assert got == {tuple(sorted(x)) for x in brute_splits(idx, size)}
    for ind0, ind1, tA, tC in sols:
        assert int(np.prod([i.size for i in ind0])) == size
        assert int(np.prod([i.size for i in ind0 + ind1])) == int(np.prod(sizes))
        assert tA.size() == A.size() and tC.size() == C.size()


def test_split_index_set_stride_one_rules():
    # config (i) m-set [a(72), b(72)] at 192 with A, C stride-one 'a' kept and
    # C's part a multiple of 8 (registerSize for float on AVX2)
    idx = [Index("a", 72), Index("b", 72)]
    A = T("A", [("a", 72), ("e", 48), ("b", 72), ("f", 48)])
    C = T("C", [("a", 72), ("b", 72), ("c", 72), ("d", 72)])
    sols = split_index_set(idx, 192, A, True, C, True, 8, True)
    assert sols
    for ind0, ind1, tA, tC in sols:
        a = [i for i in ind0 if i.label.startswith("a")]
        assert a and a[0].size % 8 == 0
        assert tA.indices[0].label.startswith("a") and tA.ld[0] == 1
    # infeasible: 'a' must be kept but 7 divides neither extent
    assert split_index_set(idx, 7, A, True, C, True) == []


def test_split_keeps_strides():
    A = T("A", [("a", 72), ("e", 48), ("b", 72)])
    C = T("C", [("a", 72), ("b", 72)])
    sols = split_index_set([Index("a", 72), Index("b", 72)], 24, A, True, C, True)
    for ind0, ind1, tA, tC in sols:
        # a split index keeps the parent stride for _0 and stride*size for _1
        for t in (tA, tC):
            assert t.size() == int(np.prod([i.size for i in t.indices]))
            a0 = t.pos(Index("a_0", 0))
            if a0 >= 0:
                assert t.ld[a0 + 1] == t.ld[a0] * t.indices[a0].size


def test_sub_tensor_and_perm():
    A = T("A", [("a", 4), ("k", 5), ("b", 6)])
    sub = A.sub_tensor([Index("b", 0), Index("a", 0)])
    assert [i.label for i in sub.indices] == ["a", "b"] and sub.ld == [1, 20]
    packed = T("At", [("b", 6), ("a", 4)])
    assert get_perm(sub, packed) == [1, 0]
    perm, size, lda, ldb = transpose_args(sub, packed)
    assert size == [4, 6] and lda == [20, 6] and ldb == [6, 4]


def test_contiguous_count():
    A = T("A", [("a", 4), ("k", 5), ("b", 6)])
    assert A.count_contiguous_stride_one_elements() == 120
    sub = A.sub_tensor([Index("a", 0), Index("b", 0)])
    assert sub.count_contiguous_stride_one_elements() == 4
