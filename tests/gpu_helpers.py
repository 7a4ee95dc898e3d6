"""Shared helpers of the GPU parity tests: seeded inputs in TCCG layout, a run
through the C ABI, and the FP64 einsum oracle."""
from __future__ import annotations

import numpy as np

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.plan import F64

TOL = {0: 1e-5, 1: 1e-13, 2: 1e-5, 3: 8e-2}


def inputs(c: Contraction, seed: int, dist: str = "uniform", dtype=np.float32, with_c: bool = False):
    """A then B (then C) drawn from default_rng(seed) in FP64 and cast."""
    rng = np.random.default_rng(seed)
    sa, sb, sc = c.shapes()
    draw = (lambda s: rng.uniform(-1.0, 1.0, s)) if dist == "uniform" else (lambda s: rng.standard_normal(s))
    A = draw(sa).astype(dtype)
    B = draw(sb).astype(dtype)
    C = draw(sc).astype(dtype) if with_c else np.zeros(sc, dtype=dtype)
    return A, B, C


def run_gpu(c: Contraction, A, B, C, prec=0, alpha=None, beta=None, tile=None, plan=None):
    import torch
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (A, B, C))
    dp = c.device_plan(prec, tile, plan=plan)
    dp.run(tA, tB, tC, c.alpha if alpha is None else alpha, c.beta if beta is None else beta)
    torch.cuda.synchronize()
    return tC.cpu().numpy()


def check(c: Contraction, prec=0, seed=0, alpha=None, beta=None, tile=None, plan=None, dist="uniform",
          tol=None):
    dtype = np.float64 if prec == F64 else np.float32
    alpha = c.alpha if alpha is None else alpha
    beta = c.beta if beta is None else beta
    A, B, C = inputs(c, seed, dist, dtype, with_c=beta != 0)
    out = run_gpu(c, A, B, C, prec, alpha, beta, tile, plan)
    ref = contract_ref([i.label for i in c.C0.indices], [i.label for i in c.A0.indices],
                       [i.label for i in c.B0.indices], A, B, C, alpha, beta)
    err = rel_frobenius(out, ref)
    limit = TOL[prec] if tol is None else tol
    assert np.isfinite(err) and err <= limit, "rel. Frobenius error %.3e > %.1e (%s, prec %d)" % (err, limit, c, prec)
    return err
