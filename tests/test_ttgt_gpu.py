"""TTGT on the GPU (transposition kernels around this backend's GETT) against
the FP64 einsum oracle; candidate bookkeeping is checked on the CPU by
tests/test_ttgt.py."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.plan import F32, F64
from tccg_amd.ttgt import DeviceTtgt, best_candidates, ttgt_candidates
from tccg_amd.workloads import WORKLOADS, scaled

pytestmark = pytest.mark.gpu


def _some(cands):
    """The model's best candidates plus one that transposes C and one that
    exchanges A and B (when they exist)."""
    out = best_candidates(cands, 4)
    for pred in (lambda c: c.transC, lambda c: c.swap == 1 and c.opA == "T"):
        extra = next((c for c in cands if pred(c)), None)
        if extra is not None and extra not in out:
            out.append(extra)
    return out


@pytest.mark.parametrize("wl", ["example", "abcd-aebf-dfce", "abc-akb-kc", "abcdef-dgfa-gceb"])
@pytest.mark.parametrize("prec", [F32, F64])
def test_ttgt_candidates(wl, prec):
    import torch
    c = Contraction.from_tccg(scaled(wl, 0.25) if wl != "example" else WORKLOADS[wl])
    dtype = np.float64 if prec == F64 else np.float32
    rng = np.random.default_rng(3)
    sa, sb, sc = c.shapes()
    A, B, C = (rng.uniform(-1, 1, s).astype(dtype) for s in (sa, sb, sc))
    labels = lambda T: [i.label for i in T.indices]
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    tol = 1e-13 if prec == F64 else 1e-5
    for cand in _some(ttgt_candidates(c.problem, 8 if prec == F64 else 4)):
        print("ttgt", wl, prec, cand.name, flush=True)
        d = DeviceTtgt(c, cand, prec)
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            tC = torch.from_numpy(C).to(dev)
            d.run(tA, tB, tC, alpha, beta)
            torch.cuda.synchronize()
            ref = contract_ref(labels(c.C0), labels(c.A0), labels(c.B0), A, B, C, alpha, beta)
            assert rel_frobenius(tC.cpu().numpy(), ref) < tol, (cand.name, alpha, beta)
