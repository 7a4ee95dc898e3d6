"""Acceptance bars of the GETT path at the full workload sizes, checked
against the FP64 einsum oracle on the same seeded inputs:

- compute-bound contraction C[a,b,c,d] = A[a,e,b,f] * B[d,f,c,e] (72/48),
  iid uniform[-1,1]: plain FP32 within 1e-5 and FP64 within 1e-13 relative
  Frobenius error;
- precision contraction C[a,b,c,d,e,f] = A[d,g,f,a] * B[g,c,e,b] (24/96),
  iid N(0,1): split-BF16 at most twice plain FP32's error, FP32 within 1e-5;
  FP8 (per-tensor scaling) reported and bounded by its quantisation error.

Each runs the plan the model ranks first (no autotuning), through the C ABI.
"""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.plan import BF16X3, F32, F64, FP8
from tccg_amd.workloads import WORKLOADS

from .gpu_helpers import inputs, run_gpu

pytestmark = pytest.mark.gpu


def _ref(c, A, B):
    return contract_ref([i.label for i in c.C0.indices], [i.label for i in c.A0.indices],
                        [i.label for i in c.B0.indices], A, B)


@pytest.mark.parametrize("prec,tol", [(F32, 1e-5), (F64, 1e-13)])
def test_compute_bound_contraction_full_size(prec, tol):
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    dtype = np.float64 if prec == F64 else np.float32
    A, B, C = inputs(c, 0, "uniform", dtype)
    print("full size", c, "prec", prec, c.plan(prec).describe(), flush=True)
    err = rel_frobenius(run_gpu(c, A, B, C, prec), _ref(c, A, B))
    print("rel. Frobenius error %.3e" % err, flush=True)
    assert np.isfinite(err) and err <= tol


def test_precision_contraction_split_bf16_within_twice_fp32_error():
    c = Contraction.from_tccg(WORKLOADS["abcdef-dgfa-gceb"])
    A, B, C = inputs(c, 0, "normal", np.float32)
    ref = _ref(c, A, B)
    errs = {}
    for prec in (F32, BF16X3, FP8):
        print("full size", c, "prec", prec, c.plan(prec).describe(), flush=True)
        errs[prec] = rel_frobenius(run_gpu(c, A, B, C, prec), ref)
        print("rel. Frobenius error %.3e" % errs[prec], flush=True)
    assert errs[F32] <= 1e-5
    assert errs[BF16X3] <= 2 * errs[F32]
    # e4m3 keeps 3 mantissa bits: ~2^-5 relative per operand, partly averaged out over k = 96
    assert errs[FP8] < 8e-2
