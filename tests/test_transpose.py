"""Tensor transposition: the oracle (pinned by brute force and closed forms),
the planner's index sets and tables, and a CPU model of both kernels'
addressing (csrc/transpose.hpp) that asserts every access is in bounds."""
import itertools

import numpy as np
import pytest

from oracle.einsum_ref import rel_frobenius, transpose_ref
from tccg_amd.transpose import INVALID, TRANSPOSE_TILES, fuse_groups, plan_transpose


def _brute(in_labels, out_labels, sizes, X):
    """Element by element through explicit column-major strides."""
    flat = np.asarray(X, dtype=np.float64).reshape(-1)
    out = np.zeros(flat.size)

    def strides(labels):
        s, o = 1, {}
        for l in labels:
            o[l] = s
            s *= sizes[l]
        return o

    si, so = strides(in_labels), strides(out_labels)
    for idx in itertools.product(*[range(sizes[l]) for l in in_labels]):
        pos = dict(zip(in_labels, idx))
        out[sum(pos[l] * so[l] for l in out_labels)] = flat[sum(pos[l] * si[l] for l in in_labels)]
    return out.reshape([sizes[l] for l in reversed(out_labels)])


def _rand(labels, sizes, seed):
    return np.random.default_rng(seed).standard_normal([sizes[l] for l in reversed(labels)])


PERMS = [("ab", "ba"), ("abc", "cab"), ("abc", "bca"), ("abc", "acb"), ("abcd", "dcba"), ("abcd", "badc"),
         ("abcd", "acbd"), ("abcde", "ecabd")]


@pytest.mark.parametrize("pi", range(len(PERMS)))
def test_oracle_matches_brute_force(pi):
    i, o = PERMS[pi]
    sizes = {l: 2 + k for k, l in enumerate(i)}
    X = _rand(i, sizes, pi)
    np.testing.assert_array_equal(transpose_ref(i, o, X), _brute(i, o, sizes, X))
    # a wrong permutation (the output order reversed) is caught by the same check
    wrong = transpose_ref(i, o[::-1], X).reshape(-1)
    assert not np.array_equal(wrong, _brute(i, o, sizes, X).reshape(-1))


def test_oracle_closed_forms():
    X = _rand("ab", dict(a=5, b=3), 1)
    np.testing.assert_array_equal(transpose_ref("ab", "ba", X), X.T)       # matrix transpose
    Y = _rand("abc", dict(a=4, b=3, c=2), 2)
    O = _rand("abc", dict(a=4, b=3, c=2), 3)
    np.testing.assert_allclose(transpose_ref("abc", "abc", Y, 2.0, -0.5, O), 2.0 * Y - 0.5 * O)   # identity
    T = transpose_ref("abc", "cab", Y)
    np.testing.assert_array_equal(transpose_ref("cab", "abc", T), Y)       # inverse permutation


def test_fuse_groups():
    assert fuse_groups("abcd", "cdab") == [["a", "b"], ["c", "d"]]
    assert fuse_groups("abcd", "dcba") == [["a"], ["b"], ["c"], ["d"]]
    assert fuse_groups("abcde", "abdec") == [["a", "b"], ["c"], ["d", "e"]]


def emulate(plan, X, OUT, alpha, beta):
    """What the kernels compute, by their own addressing."""
    src = np.asarray(X, dtype=np.float64).reshape(-1)
    dst = np.asarray(OUT, dtype=np.float64).reshape(-1).copy()
    n = src.size
    inZ, outZ = plan.inZ.astype(np.int64), plan.outZ.astype(np.int64)
    assert plan.nZ >= 1
    if plan.tile == 0:
        L = plan.Xp
        assert L * plan.nZ == n and plan.Yp == 1
        if plan.vec > 1:
            assert L % plan.vec == 0 and np.all(inZ % plan.vec == 0) and np.all(outZ % plan.vec == 0)
        assert inZ.max() + L <= n and outZ.max() + L <= n
        idx_in = (inZ[:, None] + np.arange(L)[None, :]).reshape(-1)
        idx_out = (outZ[:, None] + np.arange(L)[None, :]).reshape(-1)
        assert np.unique(idx_out).size == n         # every output element written once
        dst[idx_out] = alpha * src[idx_in] + (beta * dst[idx_out] if beta != 0 else 0)
        return dst.reshape(np.shape(OUT))
    TX, TY = TRANSPOSE_TILES[plan.tile]
    assert plan.Xp % TX == 0 and plan.Yp % TY == 0
    vx = plan.inX != INVALID
    vy = plan.inY != INVALID
    assert np.array_equal(vx, plan.outX != INVALID) and np.array_equal(vy, plan.outY != INVALID)
    iX, oX = plan.inX[vx].astype(np.int64), plan.outX[vx].astype(np.int64)
    iY, oY = plan.inY[vy].astype(np.int64), plan.outY[vy].astype(np.int64)
    # consecutive x read consecutive input elements, consecutive y write consecutive output elements
    assert np.all(np.diff(iX) == 1) and np.all(np.diff(oY) == 1)
    if plan.vec > 1:
        # groups of vec positions: whole, and every access starts on a multiple of vec
        v = plan.vec
        assert iX.size % v == 0 and oY.size % v == 0
        for a in (iY, inZ, oX, outZ, iX[::v], oY[::v]):
            assert np.all(a % v == 0)
    assert inZ.max() + iX.max() + iY.max() < n and outZ.max() + oX.max() + oY.max() < n
    idx_in = (inZ[:, None, None] + iY[None, :, None] + iX[None, None, :]).reshape(-1)
    idx_out = (outZ[:, None, None] + oY[None, :, None] + oX[None, None, :]).reshape(-1)
    assert idx_in.size == n and np.unique(idx_out).size == n
    dst[idx_out] = alpha * src[idx_in] + (beta * dst[idx_out] if beta != 0 else 0)
    return dst.reshape(np.shape(OUT))


CASES = [("ab", "ba", dict(a=70, b=33)), ("abc", "cab", dict(a=24, b=5, c=40)), ("abc", "acb", dict(a=12, b=9, c=7)),
         ("abcd", "dcba", dict(a=5, b=7, c=3, d=4)), ("abcd", "badc", dict(a=8, b=16, c=3, d=5)),
         ("abcd", "cdab", dict(a=6, b=10, c=9, d=4)), ("abcdef", "dbfaec", dict(a=4, b=3, c=5, d=2, e=3, f=4)),
         ("abcd", "abdc", dict(a=8, b=2, c=3, d=5))]


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_plan_model_matches_oracle(ci):
    i, o, sizes = CASES[ci]
    X = _rand(i, sizes, ci)
    OUT = _rand(o, sizes, ci + 50)
    ref = transpose_ref(i, o, X, 1.5, -0.25, OUT)
    tiles = [0] if fuse_groups(i, o)[0][0] == o[0] else range(1, len(TRANSPOSE_TILES))
    for t in tiles:
        plan = plan_transpose(i, o, sizes, tile=t)
        xl = [l for g in plan.X for l in g]
        yl = [l for g in plan.Y for l in g]
        zl = [l for g in plan.Z for l in g]
        assert sorted(xl + yl + zl) == sorted(i)
        assert list(i[:len(xl)]) == xl and (not yl or list(o[:len(yl)]) == yl)
        np.testing.assert_array_equal(emulate(plan, X, OUT, 1.0, 0.0).reshape(-1), transpose_ref(i, o, X).reshape(-1))
        assert rel_frobenius(emulate(plan, X, OUT, 1.5, -0.25), ref) < 1e-15


def test_tile_vector_width():
    assert plan_transpose("abc", "cab", dict(a=24, b=5, c=40)).vec == 4
    assert plan_transpose("abc", "cab", dict(a=24, b=5, c=42), elem_bytes=8).vec == 2
    assert plan_transpose("abc", "cab", dict(a=24, b=5, c=42)).vec == 1
    assert plan_transpose("abcd", "dcba", dict(a=5, b=7, c=3, d=4)).vec == 1


def test_row_copy_vector_width():
    p = plan_transpose("abcd", "abdc", dict(a=8, b=2, c=3, d=5), elem_bytes=4)
    assert p.tile == 0 and p.vec == 4 and p.Xp == 16
    p = plan_transpose("abc", "acb", dict(a=6, b=3, c=5), elem_bytes=4)
    assert p.tile == 0 and p.vec == 1
    p = plan_transpose("abc", "acb", dict(a=6, b=3, c=5), elem_bytes=8)
    assert p.tile == 0 and p.vec == 2
    with pytest.raises(ValueError):
        plan_transpose("ab", "ba", dict(a=4, b=4), tile=0)
    with pytest.raises(ValueError):
        plan_transpose("ab", "ac", dict(a=4, b=4, c=4))
