"""LoG on the GPU (the batched GETT launch over shared GEMM tables) against
the FP64 einsum oracle; each case's addressing is modelled on the CPU by
tests/test_loop_gemm.py."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.loop_gemm import DeviceLoG, log_candidates
from tccg_amd.workloads import scaled

from .test_loop_gemm import CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wl,f", CASES)
def test_log_candidates(wl, f):
    import torch
    c = Contraction.from_tccg(scaled(wl, f))
    rng = np.random.default_rng(5)
    sa, sb, sc = c.shapes()
    A, B, C = (rng.uniform(-1, 1, s).astype(np.float32) for s in (sa, sb, sc))
    lab = lambda T: [i.label for i in T.indices]
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    for cand in log_candidates(c.problem)[:6]:
        for alpha, beta in ((1.0, 0.0), (1.5, -0.5)):
            print("log", wl, cand.name, alpha, beta, flush=True)
            tC = torch.from_numpy(C).to(dev)
            DeviceLoG(c, cand, tA, tB, tC, alpha, beta).launch()
            torch.cuda.synchronize()
            ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, alpha, beta)
            assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5, (cand.name, alpha, beta)
