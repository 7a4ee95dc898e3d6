"""LoG on the GPU (every GEMM instance a problem of one grouped packed launch,
contracted loops folded into k) against the FP64 einsum oracle in FP32, FP64
and split-BF16; each case's addressing is modelled on the CPU by
tests/test_loop_gemm.py."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.loop_gemm import DeviceLoG, log_candidates
from tccg_amd.plan import BF16X3, F64, PREC_NAMES
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
            d = DeviceLoG(c, cand, tA, tB, tC, alpha, beta)
            assert d.nlaunches == 2
            d.launch()
            torch.cuda.synchronize()
            ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, alpha, beta)
            assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5, (cand.name, alpha, beta)


@pytest.mark.parametrize("prec,tol", [(F64, 1e-13), (BF16X3, 1e-5)])
@pytest.mark.parametrize("wl,f", [("abcd-aebf-dfce", 0.34), ("abcd-aklb-lckd", 0.12), ("abc-akb-kc", 0.06)])
def test_log_precisions(prec, tol, wl, f):
    import torch
    c = Contraction.from_tccg(scaled(wl, f))
    rng = np.random.default_rng(6)
    dt = np.float64 if prec == F64 else np.float32
    A, B, C = (rng.uniform(-1, 1, s).astype(dt) for s in c.shapes())
    lab = lambda T: [i.label for i in T.indices]
    dev = torch.device("cuda", 0)
    tA, tB = torch.from_numpy(A).to(dev), torch.from_numpy(B).to(dev)
    for cand in log_candidates(c.problem)[:2]:
        print("log", PREC_NAMES[prec], wl, cand.name, flush=True)
        tC = torch.from_numpy(C).to(dev)
        DeviceLoG(c, cand, tA, tB, tC, 0.75, 1.25, prec=prec).launch()
        torch.cuda.synchronize()
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, 0.75, 1.25)
        assert rel_frobenius(tC.cpu().numpy(), ref) < tol, (cand.name, prec)


def test_log_full_size_fp64():
    """The compute-bound contraction (72/48) by LoG in FP64, the model's best
    candidate, against FP64 torch.einsum on the GPU (both FP64: the bar is
    the FP64 GETT's 1e-13)."""
    import string
    import torch
    from tccg_amd.loop_gemm import best_log_candidates
    from tccg_amd.workloads import WORKLOADS
    c = Contraction.from_tccg(WORKLOADS["abcd-aebf-dfce"])
    cand = best_log_candidates(log_candidates(c.problem), 1)[0]
    g = torch.Generator(device="cuda").manual_seed(0)
    A = torch.rand(c.A0.torch_shape(), device="cuda", dtype=torch.float64, generator=g) * 2 - 1
    B = torch.rand(c.B0.torch_shape(), device="cuda", dtype=torch.float64, generator=g) * 2 - 1
    C = torch.zeros(c.C0.torch_shape(), device="cuda", dtype=torch.float64)
    DeviceLoG(c, cand, A, B, C, 1.0, 0.0, prec=F64).launch()
    torch.cuda.synchronize()
    lab = lambda T: [i.label for i in T.indices]
    letters = dict(zip(sorted({l for T in (c.C0, c.A0, c.B0) for l in lab(T)}), string.ascii_letters))
    rev = lambda T: "".join(letters[l] for l in reversed(lab(T)))
    ref = torch.einsum("%s,%s->%s" % (rev(c.A0), rev(c.B0), rev(c.C0)), A, B)
    err = float((C - ref).norm() / ref.norm())
    assert err < 1e-13, err
