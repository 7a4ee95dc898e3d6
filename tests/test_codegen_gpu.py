"""Generated per-contraction libraries (the reference's gett(A,B,C,alpha,beta)
entry point) against the FP64 einsum oracle on the GPU."""
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction, codegen
from tccg_amd.plan import F32
from tccg_amd.runtime import GeneratedGett
from tccg_amd.workloads import WORKLOADS

from .gpu_helpers import inputs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("wl", ["example", "cpu48", "abcd-aebf-dfce", "abc-akb-kc"])
def test_generated_workload_library(wl):
    import torch
    lib_path = codegen.GENERATED_DIR / ("lib%s.so" % codegen.workload_library_name(wl, F32))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(WORKLOADS[wl])
    assert (g.info["M"], g.info["N"], g.info["K"]) == (c.problem.M, c.problem.N, c.problem.K)
    A, B, C = inputs(c, seed=11)
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    g(tA, tB, tC, 1.0, 0.0)
    torch.cuda.synchronize()
    ref = contract_ref([i.label for i in c.C0.indices], [i.label for i in c.A0.indices],
                       [i.label for i in c.B0.indices], A, B)
    assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5
    # beta path through the generated entry point: C <- 2*AB + 0.5*C
    tC2 = torch.from_numpy(np.ascontiguousarray(ref.astype(np.float32))).to(dev)
    g(tA, tB, tC2, 2.0, 0.5)
    torch.cuda.synchronize()
    assert rel_frobenius(tC2.cpu().numpy(), 2.5 * ref) < 1e-5


def test_generated_library_vector_operand_a():
    """C[j,i] = 1.5 * A[k,h] * B[h,i,k,j]: A holds no index of C, so the
    generated gett() passes B as the internal A operand (empty n set)."""
    import torch
    wl = "ji-kh-hikj"
    lib_path = codegen.GENERATED_DIR / ("lib%s.so" % codegen.workload_library_name(wl, F32))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(codegen.VECTOR_A_TEXTS[wl])
    assert c.problem.swapped and c.problem.N == 1
    assert (g.info["M"], g.info["N"], g.info["K"]) == (c.problem.M, c.problem.N, c.problem.K)
    A, B, C = inputs(c, seed=4, dist="normal", with_c=True)
    dev = torch.device("cuda", 0)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for alpha, beta in ((1.5, 0.0), (-0.5, 2.0)):
        tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
        g(tA, tB, tC, alpha, beta)
        torch.cuda.synchronize()
        ref = contract_ref(*labels, A, B, C, alpha, beta)
        assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5, (alpha, beta)


@pytest.mark.parametrize("wl", sorted(codegen.EDGE_TEXTS))
def test_generated_library_edge_index_sets(wl):
    """Generated libraries of contractions at the edges of the index sets
    (empty n or k set, extent-1 index, A without a free index, fused indices
    with exchanged operands), each with the planner's best plan, through the
    generated gett() with beta == 0 and beta != 0."""
    import torch
    lib_path = codegen.GENERATED_DIR / ("lib%s.so" % codegen.workload_library_name(wl, F32))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(codegen.EDGE_TEXTS[wl])
    assert (g.info["M"], g.info["N"], g.info["K"]) == (c.problem.M, c.problem.N, c.problem.K)
    A, B, C = inputs(c, seed=8, dist="normal", with_c=True)
    dev = torch.device("cuda", 0)
    labels = [[i.label for i in T.indices] for T in (c.C0, c.A0, c.B0)]
    for alpha, beta in ((1.0, 0.0), (0.75, -1.5)):
        print("generated edge", wl, g.info, alpha, beta, flush=True)
        tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
        g(tA, tB, tC, alpha, beta)
        torch.cuda.synchronize()
        ref = contract_ref(*labels, A, B, C, alpha, beta)
        assert rel_frobenius(tC.cpu().numpy(), ref) < 1e-5, (alpha, beta)


def test_cli_testing_on_gpu(tmp_path):
    f = tmp_path / "transC.tccg"
    f.write_text("C[i1,j,i2] = 2.4 * A[i1,k,i2] * B[j,k] + 1.1 * C[i1,j,i2]\ni1 = 96\ni2 = 48\nj = 64\nk = 8\n")
    env = dict(os.environ, TCCG_DB=str(tmp_path / "t.db"), PYTHONPATH=ROOT)
    proc = subprocess.run([sys.executable, "-m", "tccg_amd", "--testing", "--maxImplementations=3",
                           "--output", str(tmp_path / "impl"), str(f)], cwd=str(tmp_path), env=env,
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
    assert proc.returncode == 0, proc.stdout
    assert "GFLOPS // estimated:" in proc.stdout and "Best Loop/TTGT/GETT/reference/GEMM/TCL:" in proc.stdout
    assert "Best TTGT-based implementation (ttgemmt_" in proc.stdout   # TTGT timed and checked too
    assert "ERROR" not in proc.stdout
    # second run reuses the database entry
    proc2 = subprocess.run([sys.executable, "-m", "tccg_amd", "--output", str(tmp_path / "impl2"), str(f)],
                           cwd=str(tmp_path), env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                           timeout=600)
    assert proc2.returncode == 0 and "[database] reusing" in proc2.stdout


@pytest.mark.parametrize("wl,prec,suffix,tol", [("abcd-aebf-dfce", 2, "_packed", 1e-5),
                                                ("abcdef-dgfa-gceb", 3, "_packed", 8e-2),
                                                ("abcdef-dgfa-gceb", 3, "_storewaves", 8e-2)])
def test_generated_packed_library(wl, prec, suffix, tol):
    """Generated libraries of packed plans (pack kernels + DMA-ring GETT kernel,
    or the store-wave kernel; workspace allocated by the library on its first
    call)."""
    import torch
    lib_path = codegen.GENERATED_DIR / ("lib%s%s.so" % (codegen.workload_library_name(wl, prec), suffix))
    assert lib_path.exists(), "generated library missing: run __graft_entry__.build()"
    g = GeneratedGett(lib_path)
    assert g.info["prec"] == prec
    c = Contraction.from_tccg(WORKLOADS[wl])
    A, B, C = inputs(c, seed=5, dist="normal")
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    ref = contract_ref([i.label for i in c.C0.indices], [i.label for i in c.A0.indices],
                       [i.label for i in c.B0.indices], A, B)
    for _ in range(2):   # the second call reuses the workspace
        tC.zero_()
        g(tA, tB, tC, 1.0, 0.0)
        torch.cuda.synchronize()
        assert rel_frobenius(tC.cpu().numpy(), ref) < tol


def test_generated_library_rejects_bad_tensors():
    """The generated entry point has its extents baked in: the wrapper checks
    dtype, element count and device before the launch."""
    import torch
    from tccg_amd.runtime import GettRuntimeError
    lib_path = codegen.GENERATED_DIR / ("lib%s.so" % codegen.workload_library_name("example", F32))
    g = GeneratedGett(lib_path)
    c = Contraction.from_tccg(WORKLOADS["example"])
    assert (g.info["sizeA"], g.info["sizeB"], g.info["sizeC"]) == (c.A0.size(), c.B0.size(), c.C0.size())
    A, B, C = inputs(c, seed=1)
    dev = torch.device("cuda", 0)
    tA, tB, tC = (torch.from_numpy(x).to(dev) for x in (A, B, C))
    with pytest.raises(TypeError):
        g(tA.double(), tB, tC)
    with pytest.raises(GettRuntimeError):
        g(tA, tB, tC.reshape(-1)[: tC.numel() - 1])
    with pytest.raises(GettRuntimeError):
        g(tA, tB, tC.cpu())
    g(tA, tB, tC)   # the well-formed call still runs
    torch.cuda.synchronize()
