"""The oracle pinned by closed forms, brute force and the golden fixtures (CPU)."""
import glob
import itertools
import os

import numpy as np
import pytest

from oracle.avx2_gett import Avx2Gett
from oracle.avx2_model import get_possible_values as possible_values
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
    # near zero the absolute difference counts (min of relative and absolute)
    assert reference_equal(np.array([0.0]), np.array([5e-5]), False)
    assert not reference_equal(np.array([0.0]), np.array([5e-4]), False)


def test_reference_inputs_known_values():
    # hand-computed: A[0] = (10 + 17) / 110, A[5] = (10 + 102 % 100) / 110,
    # B[0] = (10 + 6) / 110, C[29] = (10 + 416 % 100) / 110
    assert reference_inputs(6, "A")[[0, 5]].tolist() == [np.float32(27 / 110.0), np.float32(12 / 110.0)]
    assert reference_inputs(1, "B")[0] == np.float32(16 / 110.0)
    assert reference_inputs(30, "C")[29] == np.float32(26 / 110.0)
    assert 0.09 <= reference_inputs(1000, "C").min() and reference_inputs(1000, "B").max() <= 1.0


def test_cli_harness_matches_the_oracle():
    """The CLI's restatement of the generated harness (inputs and element
    check, tccg_amd/cli.py) agrees with the oracle's on the CPU."""
    import torch
    from tccg_amd.cli import harness_equal, harness_inputs
    for which in "ABC":
        for dt, ndt in ((torch.float32, np.float32), (torch.float64, np.float64)):
            got = harness_inputs(5000, which, dt, device="cpu").numpy()
            assert np.array_equal(got, reference_inputs(5000, which, ndt)), (which, dt)
    rng = np.random.default_rng(4)
    for scale in (1e-6, 5e-5, 2e-4, 1e-3):
        x = rng.uniform(-1, 1, 4096)
        y = x * (1 + scale * rng.standard_normal(4096))
        for dbl in (False, True):
            assert harness_equal(torch.from_numpy(y), torch.from_numpy(x), dbl) == reference_equal(y, x, dbl)


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
    # config (i): the restated planner's best-ranked candidate is the m-outer
    # variant with mc from getPossibleValues(24, 5184, 4600, 3), nc from
    # getPossibleValues(4, 5184, 192, 3), kc from getPossibleValues(1, 2304, 256, 4)
    g = Avx2Gett(["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"],
                 dict(a=72, b=72, c=72, d=72, e=48, f=48))
    assert (g.variant, g.mc, g.nc, g.kc) == (1, 2592, 192, 256)
    assert g.mc == possible_values(24, 5184, 4600, 3)[0] and g.nc == possible_values(4, 5184, 192, 3)[0]
    assert g.kc == possible_values(1, 2304, 256, 4)[0]


@pytest.mark.parametrize("variant", [0, 1])
def test_avx2_gett_both_variants(variant):
    from oracle.avx2_model import candidates_for
    c, a, b = ["a", "b", "c"], ["c", "k", "a"], ["k", "b"]
    sizes = dict(a=48, b=24, c=20, k=96)
    cands = [x for x in candidates_for(c, a, b, sizes, "float", 0.0, -1) if x.variant == variant]
    assert cands
    rng = np.random.default_rng(variant)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(np.float32)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(np.float32)
    C0 = rng.uniform(-1, 1, [sizes[l] for l in reversed(c)]).astype(np.float32)
    ref = contract_ref(c, a, b, A, B, C0, 0.5, 2.0)
    for cand in cands[:4]:
        g = Avx2Gett(c, a, b, sizes, candidate=cand)
        C = C0.copy()
        g(A, B, C, 0.5, 2.0, threads=3)
        assert rel_frobenius(C, ref) < 1e-5, cand.name()
    assert not Avx2Gett(["a", "b"], ["a", "k"], ["k", "b"], dict(a=20, b=8, k=4)).supported()
