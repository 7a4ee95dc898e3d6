"""Differentiable GETT: gradcheck in FP64, finite differences, torch.einsum autograd."""
import numpy as np
import pytest

from tccg_amd import Contraction
from tccg_amd.autograd import contract
from tccg_amd.workloads import WORKLOADS, scaled

pytestmark = pytest.mark.gpu

PATTERNS = [
    (["a", "b", "c", "d"], ["a", "k", "l", "b"], ["l", "c", "k", "d"]),
    (["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"]),
    (["a", "b", "c"], ["c", "k", "a"], ["k", "b"]),
    (["b", "a"], ["a", "k"], ["k", "b"]),
]


def einsum_eq(c):
    import string
    letters, pool = {}, iter(string.ascii_letters)
    for T in (c.C0, c.A0, c.B0):
        for i in T.indices:
            letters.setdefault(i.label, next(pool))
    rev = lambda T: "".join(letters[i.label] for i in reversed(T.indices))
    return "%s,%s->%s" % (rev(c.A0), rev(c.B0), rev(c.C0))


@pytest.mark.parametrize("pi", range(len(PATTERNS)))
def test_gradcheck_fp64_extents_6(pi):
    import torch
    c_l, a_l, b_l = PATTERNS[pi]
    sizes = {l: 6 for l in set(c_l + a_l + b_l)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes, alpha=1.3)
    g = torch.Generator(device="cuda").manual_seed(pi)
    A = torch.rand(c.A0.torch_shape(), dtype=torch.float64, device="cuda", generator=g, requires_grad=True)
    B = torch.rand(c.B0.torch_shape(), dtype=torch.float64, device="cuda", generator=g, requires_grad=True)
    assert torch.autograd.gradcheck(lambda x, y: contract(x, y, c, "f64"), (A, B), eps=1e-6, atol=1e-8, rtol=1e-6)


def test_central_differences_fp64():
    import torch
    c_l, a_l, b_l = PATTERNS[0]
    sizes = {l: 6 for l in set(c_l + a_l + b_l)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes)
    gen = torch.Generator(device="cuda").manual_seed(5)
    A = torch.randn(c.A0.torch_shape(), dtype=torch.float64, device="cuda", generator=gen, requires_grad=True)
    B = torch.randn(c.B0.torch_shape(), dtype=torch.float64, device="cuda", generator=gen)
    T = torch.randn(c.C0.torch_shape(), dtype=torch.float64, device="cuda", generator=gen)
    loss = lambda X: 0.5 * ((contract(X, B, c, "f64") - T) ** 2).sum()
    L = loss(A)
    L.backward()
    h = 1e-5
    flat = A.detach().reshape(-1)
    rng = np.random.default_rng(0)
    for idx in rng.choice(flat.numel(), 12, replace=False):
        e = torch.zeros_like(flat)
        e[idx] = h
        with torch.no_grad():
            fd = (loss((flat + e).reshape(A.shape)) - loss((flat - e).reshape(A.shape))) / (2 * h)
        assert abs(fd.item() - A.grad.reshape(-1)[idx].item()) <= 1e-6 * max(1.0, abs(fd.item()))


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16x3", 1e-5), ("bf16", 2e-2)])
def test_full_size_gradients_vs_torch(prec, tol):
    """Config C[a,b,c,d] = A[a,k,l,b] * B[l,c,k,d], 72/32, loss 0.5||C - T||^2."""
    import torch
    c = Contraction.from_tccg(WORKLOADS["abcd-aklb-lckd"])
    gen = torch.Generator(device="cuda").manual_seed(0)
    A = torch.randn(c.A0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    B = torch.randn(c.B0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    T = torch.randn(c.C0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    A1, B1 = A.clone().requires_grad_(), B.clone().requires_grad_()
    L = 0.5 * ((contract(A1, B1, c, prec) - T) ** 2).sum()
    L.backward()
    # FP64 torch.einsum autograd reference over contiguous operands
    A2, B2 = A.double().requires_grad_(), B.double().requires_grad_()
    L2 = 0.5 * ((torch.einsum(einsum_eq(c), A2, B2) - T.double()) ** 2).sum()
    L2.backward()
    for g1, g2 in ((A1.grad, A2.grad), (B1.grad, B2.grad)):
        err = ((g1.double() - g2).norm() / g2.norm()).item()
        assert err < tol, "%s gradient rel. error %.3e" % (prec, err)


def test_tuned_contract_gradients():
    """tune=True: forward and both adjoints get the autotuner's fastest plans;
    gradients stay those of the FP64 reference."""
    import torch
    from tccg_amd.autograd import adjoints
    c = Contraction.from_tccg(scaled("abcd-aklb-lckd", 0.25))
    gen = torch.Generator(device="cuda").manual_seed(1)
    A = torch.randn(c.A0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    B = torch.randn(c.B0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    T = torch.randn(c.C0.torch_shape(), dtype=torch.float32, device="cuda", generator=gen)
    A1, B1 = A.clone().requires_grad_(), B.clone().requires_grad_()
    L = 0.5 * ((contract(A1, B1, c, "f32", tune=True) - T) ** 2).sum()
    L.backward()
    assert all(0 in x._tuned for x in (c,) + adjoints(c))
    A2, B2 = A.double().requires_grad_(), B.double().requires_grad_()
    L2 = 0.5 * ((torch.einsum(einsum_eq(c), A2, B2) - T.double()) ** 2).sum()
    L2.backward()
    for g1, g2 in ((A1.grad, A2.grad), (B1.grad, B2.grad)):
        assert ((g1.double() - g2).norm() / g2.norm()).item() < 1e-5


def test_gradients_of_permuted_views():
    """Inputs that are permuted (non-contiguous) views: the forward contracts
    their contiguous copies and backward returns gradients of the views' shapes."""
    import torch
    c_l, a_l, b_l = PATTERNS[0]
    sizes = {l: 5 for l in set(c_l + a_l + b_l)}
    c = Contraction.from_labels(c_l, a_l, b_l, sizes, alpha=0.7)
    g = torch.Generator(device="cuda").manual_seed(3)
    sa, sb = c.A0.torch_shape(), c.B0.torch_shape()
    A0 = torch.rand(tuple(reversed(sa)), dtype=torch.float64, device="cuda", generator=g)
    B0 = torch.rand(tuple(reversed(sb)), dtype=torch.float64, device="cuda", generator=g)
    A = A0.permute(*reversed(range(A0.dim()))).requires_grad_(False)
    B = B0.permute(*reversed(range(B0.dim())))
    assert not A.is_contiguous() and A.shape == sa
    A.requires_grad_(True)
    B.requires_grad_(True)
    assert torch.autograd.gradcheck(lambda x, y: contract(x, y, c, "f64"), (A, B), eps=1e-6, atol=1e-8, rtol=1e-6)
    L = contract(A, B, c, "f64").sum()
    L.backward()
    A2 = A.detach().clone().requires_grad_()
    B2 = B.detach().clone().requires_grad_()
    torch.einsum(einsum_eq(c), A2, B2).mul(0.7).sum().backward()
    assert torch.allclose(A.grad, A2.grad, rtol=1e-12, atol=1e-12)
    assert torch.allclose(B.grad, B2.grad, rtol=1e-12, atol=1e-12)
