"""Emitted TTGT and LoG libraries (tccg_amd/codegen_methods.py, built by
__graft_entry__.build()) against the FP64 einsum oracle on the GPU: through
the uniform gett_stream() boundary with alpha/beta and a second set of
tensors (the LoG library rebuilds its records for new pointers), and
through the reference's method entry point."""
import ctypes

import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction, codegen
from tccg_amd.codegen_methods import method_library_specs
from tccg_amd.plan import F64

pytestmark = pytest.mark.gpu

SPECS = method_library_specs()


def lab(T):
    return [i.label for i in T.indices]


@pytest.mark.parametrize("spec", SPECS, ids=[s[0] for s in SPECS])
def test_emitted_method_library(spec):
    import torch
    from tccg_amd.runtime import GeneratedGett
    dname, text, kind, prec = spec
    base = "ttgemmt" if kind == "ttgt" else "loopOverGemm"
    path = codegen.GENERATED_DIR / "methods" / dname / ("lib%s.so" % base)
    assert path.exists(), "emitted library missing: run __graft_entry__.build()"
    g = GeneratedGett(path)
    c = Contraction.from_tccg(text)
    assert (g.info["M"], g.info["N"], g.info["K"]) == (c.problem.M, c.problem.N, c.problem.K)
    dt = np.float64 if prec == F64 else np.float32
    tol = 1e-13 if prec == F64 else 1e-5
    dev = torch.device("cuda", 0)
    for seed, (alpha, beta) in enumerate(((1.0, 0.0), (c.alpha if c.alpha != 1.0 else 2.5, 0.75))):
        rng = np.random.default_rng(seed)
        A, B, C = (rng.uniform(-1, 1, s).astype(dt) for s in c.shapes())
        tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
        print(dname, "call", seed, alpha, beta, flush=True)
        g(tA, tB, tC, alpha, beta)
        torch.cuda.synchronize()
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, alpha, beta)
        assert rel_frobenius(tC.cpu().numpy(), ref) < tol, (dname, seed)
    # the reference's method entry point (null stream, returns 0)
    T = ctypes.c_double if prec == F64 else ctypes.c_float
    header = (path.parent / (base + ".hpp")).read_text()
    entry = [l for l in header.splitlines() if l.startswith("int ") and "gett" not in l.split("(")[0]][0]
    fname = entry.split("(")[0].split()[-1]
    fn = getattr(g.lib, fname)
    fn.restype = ctypes.c_int
    nargs = entry.count(",") + 1
    fn.argtypes = [ctypes.c_void_p] * 3 + [T, T] + ([ctypes.c_void_p] if nargs == 6 else [])
    tC.zero_()
    torch.cuda.synchronize()
    args = [tA.data_ptr(), tB.data_ptr(), tC.data_ptr(), 1.0, 0.0] + ([None] if nargs == 6 else [])
    assert fn(*args) == 0, g.lib.gett_last_error()
    torch.cuda.synchronize()
    ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, None, 1.0, 0.0)
    assert rel_frobenius(tC.cpu().numpy(), ref) < tol, (dname, "method entry")
