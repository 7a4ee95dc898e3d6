"""The CPU emulations of the reduced-precision forms (oracle/precision_ref.py),
pinned by closed forms, hand-computed roundings and exactness invariants."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from oracle.precision_ref import (E4M3_MAX, fp8_contract, fp8_descale, fp8_quantize, fp8_scale, split_bf16_contract,
                                  split_bf16_parts)

PAT = (["a", "b"], ["a", "k"], ["k", "b"])


def test_split_parts_reconstruct_fp32_exactly():
    # (FP32 values above the largest BF16, 3.39e38, would round to infinity; within
    # 2^16 of the smallest normal the parts become subnormal)
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(65536), rng.uniform(-1e-30, 1e30, 1024),
                        [0.0, -0.0, 1.0, -3.0, 1e-33, 3.0e38]]).astype(np.float32)
    p0, p1, p2 = split_bf16_parts(x)
    assert np.array_equal(p0 + p1 + p2, x.astype(np.float64))
    import torch
    assert np.array_equal(p0, torch.from_numpy(x).to(torch.bfloat16).double().numpy())
    # every part is a BF16 value: 8 significant bits
    for p in (p0, p1, p2):
        assert np.array_equal(torch.from_numpy(p.astype(np.float32)).to(torch.bfloat16).double().numpy(), p)


@pytest.mark.parametrize("x,parts", [
    (1 + 2.0**-9 + 2.0**-18, (1.0, 2.0**-9, 2.0**-18)),
    (1 + 2.0**-8, (1.0, 2.0**-8, 0.0)),                       # a tie: rounds to the even 1.0
    (1 + 3 * 2.0**-8, (1 + 2.0**-6, -(2.0**-8), 0.0)),         # a tie: rounds up to the even neighbour
    (-(2.0 + 2.0**-8 + 2.0**-17), (-2.0, -(2.0**-8), -(2.0**-17))),   # BF16 ulp at 2 is 2^-6
    (2.0 + 2.0**-6 + 2.0**-15, (2.0 + 2.0**-6, 2.0**-15, 0.0)),
    (1 - 2.0**-24, (1.0, -(2.0**-24), 0.0)),
])
def test_split_parts_hand_computed(x, parts):
    p = split_bf16_parts(np.array([x], dtype=np.float32))
    assert tuple(float(q[0]) for q in p) == parts


def test_split_contract_keeps_six_products():
    """Two-part inputs (x2 = 0) lose nothing: the result is the exact
    contraction.  Three-part inputs lose a1 b2 + a2 b1 + a2 b2 only: the
    difference from the exact result is below 3 * 2^-24 of |A||B| per term,
    and dropping any kept product (or keeping a2 b2 instead of a2 b0) breaks
    that bound."""
    rng = np.random.default_rng(1)
    k = 16
    # two parts: 16-bit significands
    A2 = (rng.integers(-2**15, 2**15, (k, 9)) * 2.0**-15).astype(np.float32)
    B2 = (rng.integers(-2**15, 2**15, (7, k)) * 2.0**-15).astype(np.float32)
    assert np.array_equal(split_bf16_contract(*PAT, A2, B2), contract_ref(*PAT, A2, B2))
    assert np.array_equal(split_bf16_contract(*PAT, A2, B2, grouped=False), contract_ref(*PAT, A2, B2))
    A = rng.standard_normal((k, 9)).astype(np.float32)
    B = rng.standard_normal((7, k)).astype(np.float32)
    exact = contract_ref(*PAT, A, B)
    bound = 3 * 2.0**-24 * contract_ref(*PAT, np.abs(A), np.abs(B))
    got = split_bf16_contract(*PAT, A, B)
    assert np.allclose(got, split_bf16_contract(*PAT, A, B, grouped=False), rtol=1e-15, atol=1e-15)
    assert np.all(np.abs(got - exact) <= bound)
    pa, pb = split_bf16_parts(A), split_bf16_parts(B)
    for drop in ((0, 2), (1, 1), (2, 0)):
        wrong = got - contract_ref(*PAT, pa[drop[1]], pb[drop[0]])
        assert not np.all(np.abs(wrong - exact) <= bound), drop
    swapped = got - contract_ref(*PAT, pa[2], pb[0]) + contract_ref(*PAT, pa[2], pb[2])
    assert not np.all(np.abs(swapped - exact) <= bound)


@pytest.mark.parametrize("x,q", [
    (1.0625, 1.0), (1.1875, 1.25), (-1.03, -1.0), (3.3, 3.25), (500.0, 448.0), (-1e9, -448.0),
    (2.0**-9, 2.0**-9),                 # smallest subnormal
    (3 * 2.0**-11, 2.0**-9),            # 0.75 of the smallest subnormal: rounds up
    (2.0**-10, 0.0),                    # half of the smallest subnormal: a tie, rounds to the even 0
    (17.0, 16.0),                       # a tie between 16 and 18: even mantissa is 16
    (19.0, 20.0),                       # a tie between 18 and 20: even mantissa is 20
    (0.0, 0.0),
])
def test_fp8_quantize_hand_computed(x, q):
    assert float(fp8_quantize(np.array([x], dtype=np.float32), np.float32(1.0))[0]) == q


def test_fp8_values_round_trip():
    # every finite positive e4m3 value: 16 subnormals/zero, normals up to 448
    vals = [m * 2.0**-9 for m in range(8)] + [(1 + m / 8) * 2.0**e for e in range(-6, 9) for m in range(8)]
    vals = np.array([v for v in vals if v <= E4M3_MAX], dtype=np.float32)
    assert np.array_equal(fp8_quantize(vals, np.float32(1)), vals.astype(np.float64))
    assert np.array_equal(fp8_quantize(-vals, np.float32(1)), -vals.astype(np.float64))


def test_fp8_scale_and_descale():
    x = np.array([0.5, -1.75, 1.0], dtype=np.float32)
    assert fp8_scale(x) == np.float32(256.0)
    assert fp8_scale(np.zeros(4, dtype=np.float32)) == np.float32(448.0)
    y = np.array([3.5], dtype=np.float32)
    assert fp8_descale(x, y) == np.float32((1.75 / 448) * (3.5 / 448))


def test_fp8_contract_of_representable_inputs_is_exact():
    """Inputs whose scaled values are e4m3 values (amax = 1.75: scale 256)
    pass through quantisation unchanged: the emulation is the contraction."""
    rng = np.random.default_rng(2)
    grid = np.array([(1 + m / 8) * 2.0**e for e in (-2, -1) for m in range(8)] + [1.5, 1.75])
    A = (rng.choice(grid, (12, 10)) * rng.choice([-1, 1], (12, 10))).astype(np.float32)
    B = (rng.choice(grid, (6, 12)) * rng.choice([-1, 1], (6, 12))).astype(np.float32)
    A[0, 0], B[0, 0] = 1.75, -1.75
    assert np.array_equal(fp8_contract(*PAT, A, B), contract_ref(*PAT, A, B))


def test_fp8_contract_error_is_quantisation_sized():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((96, 40)).astype(np.float32)
    B = rng.standard_normal((30, 96)).astype(np.float32)
    err = rel_frobenius(fp8_contract(*PAT, A, B), contract_ref(*PAT, A, B))
    assert 1e-3 < err < 6e-2   # e4m3: 3 mantissa bits
