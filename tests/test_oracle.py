"""The oracle pinned by closed forms, brute force and the golden fixtures (CPU)."""
import glob
import itertools
import os

import numpy as np
import pytest

from oracle.avx2_gett import Avx2Gett, possible_values
from oracle.einsum_ref import contract_ref, reference_equal, reference_inputs, rel_frobenius, subscripts
from oracle.naive import naive_contract

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load_golden(path):
    z = np.load(path, allow_pickle=False)
    sizes = dict(zip([str(x) for x in z["size_labels"]], [int(x) for x in z["size_values"]]))
    return ([str(x) for x in z["c"]], [str(x) for x in z["a"]], [str(x) for x in z["b"]], sizes, z["A"], z["B"],
            z["C_in"], float(z["alpha"]), float(z["beta"]), z["C_out"])


def test_golden_present():
    assert len(GOLDEN) >= 7


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_einsum_oracle_reproduces_golden(path):
    c, a, b, sizes, A, B, C, alpha, beta, out = load_golden(path)
    ref = contract_ref(c, a, b, A, B, C, alpha, beta)
    assert np.max(np.abs(ref - out)) < 1e-13


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p) for p in GOLDEN])
def test_naive_restatement_matches_golden(path):
    c, a, b, sizes, A, B, C, alpha, beta, out = load_golden(path)
    C64 = C.astype(np.float64).copy()
    naive_contract(c, a, b, sizes, A.astype(np.float64), B.astype(np.float64), C64, alpha, beta)
    assert np.max(np.abs(C64 - out)) < 1e-12
    C32 = C.copy()
    naive_contract(c, a, b, sizes, A, B, C32, alpha, beta)
    assert rel_frobenius(C32, out) < 1e-6


def test_einsum_is_matmul_for_gemm():
    # C[i,j] = A[i,k] B[k,j] in TCCG layout: torch arrays are the transposes
    rng = np.random.default_rng(0)
    A = rng.standard_normal((7, 5))   # (k, i)
    B = rng.standard_normal((4, 7))   # (j, k)
    C = contract_ref(["i", "j"], ["i", "k"], ["k", "j"], A, B)  # (j, i)
    np.testing.assert_allclose(C.T, A.T @ B.T, rtol=1e-13)


def test_einsum_outer_product_and_kron():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(5)
    b = rng.standard_normal(3)
    C = contract_ref(["i", "j"], ["i"], ["j"], a, b)  # shape (j, i)
    np.testing.assert_allclose(C, np.outer(b, a), rtol=1e-14)


def test_einsum_exact_small_integers():
    rng = np.random.default_rng(2)
    A = rng.integers(-3, 4, (4, 6, 5)).astype(np.float32)
    B = rng.integers(-3, 4, (3, 6)).astype(np.float32)
    C = contract_ref(["a", "b", "c"], ["a", "k", "b"], ["k", "c"], A, B)
    # exact: every partial sum is a small integer
    assert np.all(C == np.round(C))
    Cb = np.zeros_like(C)
    for ia, ib, ic, ik in itertools.product(range(5), range(4), range(3), range(6)):
        Cb[ic, ib, ia] += A[ib, ik, ia] * B[ic, ik]
    assert np.array_equal(C, Cb)


def test_subscripts_reversal():
    assert subscripts(["a", "b"], ["a", "k"], ["k", "b"]) == "ca,bc->ba"


def test_planted_mistakes_fail():
    """A transposed operand, a dropped beta term or a sign flip each fail the check."""
    c, a, b, sizes, A, B, C, alpha, beta, out = load_golden(
        [p for p in GOLDEN if "transC" in p][0])
    good = contract_ref(c, a, b, A, B, C, alpha, beta)
    assert rel_frobenius(good, out) < 1e-14
    wrong_beta = contract_ref(c, a, b, A, B, C, alpha, 0.0)
    assert rel_frobenius(wrong_beta, out) > 1e-2
    wrong_sign = contract_ref(c, a, b, A, B, C, -alpha, beta)
    assert rel_frobenius(wrong_sign, out) > 1e-2
    # B's bytes read with its two indices in the wrong order (a transposed operand)
    wrong_layout = contract_ref(c, a, list(reversed(b)), A, B.reshape(B.shape[::-1]), C, alpha, beta)
    assert rel_frobenius(wrong_layout, out) > 1e-2


def test_reference_equal_semantics():
    x = np.array([1.0, 2.0, 3.0], dtype=np.float32)
    assert reference_equal(x, x * (1 + 5e-5), False)
    assert not reference_equal(x, x * (1 + 5e-4), False)
    # the reference compares absolute values: a sign flip passes its check
    assert reference_equal(x, -x, False)
    assert reference_equal(x.astype(np.float64), x.astype(np.float64) * (1 + 5e-10), True)


def test_reference_inputs_formula():
    A = reference_inputs(5, "A")
    np.testing.assert_allclose(A, [(10 + ((i + 1) * 17) % 100) / 110.0 for i in range(5)], rtol=1e-7)
    assert 0.09 <= reference_inputs(1000, "C").min() and reference_inputs(1000, "B").max() <= 1.0


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("case", [
    (["a", "b", "c", "d"], ["a", "k", "c"], ["k", "d", "b"], dict(a=48, b=8, c=24, d=12, k=64)),
    (["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"], dict(a=24, b=24, c=24, d=24, e=16, f=12)),
    (["b", "a"], ["a", "k"], ["k", "b"], dict(a=48, b=24, k=100)),
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=48, b=10, c=24, k=8)),
])
def test_avx2_gett_restatement(dt, case):
    c, a, b, sizes = case
    g = Avx2Gett(c, a, b, sizes, dt)
    assert g.supported()
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(dt)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(dt)
    C = rng.uniform(-1, 1, [sizes[l] for l in reversed(c)]).astype(dt)
    ref = contract_ref(c, a, b, A, B, C, 1.5, 0.7)
    g(A, B, C, 1.5, 0.7, threads=2)
    assert rel_frobenius(C, ref) < (1e-13 if dt == np.float64 else 1e-5)


def test_avx2_blocking_follows_reference_policy():
    # config (i): M=N=5184 (mc from {192,144,96}, nc = N < 4600? no -> 4600-ish), K=2304 (kc 256)
    g = Avx2Gett(["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"],
                 dict(a=72, b=72, c=72, d=72, e=48, f=48))
    assert g.mc == 192 and g.kc == 256 and g.nc == possible_values(4, 5184, 4600, 3)[0] == 2592
    assert not Avx2Gett(["a", "b"], ["a", "k"], ["k", "b"], dict(a=20, b=8, k=4)).supported()
