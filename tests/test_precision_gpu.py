"""The reduced-precision forms against their CPU emulations (oracle/precision_ref.py).

Bit-exact cases use inputs on which FP32 accumulation is exact in any order,
so the GPU result must equal the emulation bit for bit and any change to the
product set, the part split, the e4m3 rounding or the scales shows:
- split-BF16: x = s * 2^E * (1 + 2^-9 + x2 * 2^-18), E in {0, -1}, x2 in {0, 1}:
  parts s*2^E, s*2^(E-9), s*x2*2^(E-18); every kept product is a power of two
  of at least 2^-20 and k <= 8 keeps every partial sum below 2^4, so sums
  need at most 24 significant bits;
- FP8: |x| in [0.5, 1.75] with max|x| = 1.75 exactly, so the scale is 2^8,
  the e4m3 values are multiples of 16 below 448, their products multiples of
  256 and k = 96 sums stay below 2^25 (17 significant bits); a quarter of the
  values sit exactly halfway between two e4m3 values (round-to-even ties).
Every candidate plan of each kernel kind runs (gather, packed, persistent,
run epilogue, store waves).  At the full precision-contraction size the GPU
results must be within 1e-6 (split-BF16: FP32 accumulation of 96-term sums)
or 1e-4 (FP8) relative Frobenius error of the emulations.  The FP8 MFMA sums
its 128 products with fewer bits than FP32 (1.5e-5 relative on these N(0,1)
inputs); a wrong rounding mode or scale moves the result by about 1e-2.
"""
import numpy as np
import pytest

from oracle.einsum_ref import rel_frobenius
from oracle.precision_ref import fp8_contract, split_bf16_contract
from tccg_amd import Contraction
from tccg_amd.plan import BF16X3, FP8, candidate_plans
from tccg_amd.workloads import WORKLOADS

from .gpu_helpers import inputs, run_gpu

pytestmark = pytest.mark.gpu

CFG3 = (["a", "b", "c", "d", "e", "f"], ["d", "g", "f", "a"], ["g", "c", "e", "b"])
GEMM3 = (["a", "b", "c"], ["a", "k", "b"], ["k", "c"])


def _labels(c):
    return [i.label for i in c.C0.indices], [i.label for i in c.A0.indices], [i.label for i in c.B0.indices]


def split_exact_inputs(shape, rng):
    E = rng.choice([0.0, -1.0], shape)
    x2 = rng.integers(0, 2, shape)
    s = rng.choice([-1.0, 1.0], shape)
    return (s * 2.0**E * (1 + 2.0**-9 + x2 * 2.0**-18)).astype(np.float32)


def fp8_exact_inputs(shape, rng):
    x = rng.uniform(0.5, 1.75, shape)
    ties = np.concatenate([0.5 * (1 + (2 * np.arange(8) + 1) / 16), 1 + (2 * np.arange(6) + 1) / 16])
    mask = rng.random(shape) < 0.25
    x = np.where(mask, rng.choice(ties, shape), x)
    x = x * rng.choice([-1.0, 1.0], shape)
    x.reshape(-1)[rng.integers(0, x.size)] = -1.75
    return x.astype(np.float32)


def _plans(c, prec):
    seen, out = set(), []
    for p in candidate_plans(c.problem, prec):
        key = p.describe()
        if key not in seen:
            seen.add(key)
            out.append(p)
    return out


@pytest.mark.parametrize("pattern,sizes", [
    (GEMM3, dict(a=40, b=24, c=72, k=8)),
    (CFG3, dict(a=6, b=6, c=6, d=6, e=6, f=6, g=8)),
    (GEMM3, dict(a=64, b=3, c=33, k=5)),
])
def test_split_bf16_bit_exact(pattern, sizes):
    c = Contraction.from_labels(*pattern, sizes)
    rng = np.random.default_rng(7)
    sa, sb, sc = c.shapes()
    A, B = split_exact_inputs(sa, rng), split_exact_inputs(sb, rng)
    want = split_bf16_contract(*_labels(c), A, B).astype(np.float32)
    for plan in _plans(c, BF16X3):
        print("split-BF16 bit-exact", c, plan.describe(), flush=True)
        got = run_gpu(c, A, B, np.zeros(sc, np.float32), BF16X3, plan=plan)
        assert np.array_equal(got, want), plan.describe()


@pytest.mark.parametrize("pattern,sizes", [
    (CFG3, dict(a=12, b=12, c=12, d=12, e=12, f=12, g=96)),
    (GEMM3, dict(a=40, b=24, c=72, k=96)),
])
def test_fp8_bit_exact(pattern, sizes):
    c = Contraction.from_labels(*pattern, sizes)
    rng = np.random.default_rng(8)
    sa, sb, sc = c.shapes()
    A, B = fp8_exact_inputs(sa, rng), fp8_exact_inputs(sb, rng)
    want = fp8_contract(*_labels(c), A, B).astype(np.float32)
    for plan in _plans(c, FP8):
        print("FP8 bit-exact", c, plan.describe(), flush=True)
        got = run_gpu(c, A, B, np.zeros(sc, np.float32), FP8, plan=plan)
        assert np.array_equal(got, want), plan.describe()


@pytest.mark.parametrize("prec", [BF16X3, FP8])
def test_precision_contraction_full_size_matches_emulation(prec):
    c = Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"])
    A, B, C = inputs(c, 0, "normal", np.float32)
    print("full size", c, "prec", prec, c.plan(prec).describe(), flush=True)
    got = run_gpu(c, A, B, C, prec)
    emul = (split_bf16_contract if prec == BF16X3 else fp8_contract)(*_labels(c), A, B)
    err = rel_frobenius(got, emul)
    print("rel. Frobenius error vs emulation %.3e" % err, flush=True)
    assert err <= (1e-6 if prec == BF16X3 else 1e-4)
