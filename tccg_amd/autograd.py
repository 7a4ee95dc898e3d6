"""Differentiable GETT: a torch.autograd.Function whose backward runs the two
adjoint contractions as GETT kernels.

For C = alpha * sum_k A * B the adjoints are again tensor contractions, which
the generator derives from the forward's index lists:
    dA[A indices] = alpha * sum dC[C indices] * B[B indices]   (summed over n)
    dB[B indices] = alpha * sum A[A indices] * dC[C indices]   (summed over m)
Each is planned like any .tccg contraction (classification, A/B swap, fusion),
so its operand permutations follow from the same rules as the forward's.
"""
from __future__ import annotations

from typing import Optional, Tuple

from .contraction import Contraction
from .plan import F64, PREC_BY_NAME


def adjoints(c: Contraction) -> Tuple[Contraction, Contraction]:
    """(dA contraction, dB contraction) of forward contraction c, cached."""
    cached = getattr(c, "_adjoints", None)
    if cached is not None:
        return cached
    a = [i.label for i in c.A0.indices]
    b = [i.label for i in c.B0.indices]
    cl = [i.label for i in c.C0.indices]
    sizes = {i.label: i.size for T in (c.A0, c.B0, c.C0) for i in T.indices}
    dA = Contraction.from_labels(a, cl, b, sizes, alpha=c.alpha, beta=0.0)
    dB = Contraction.from_labels(b, a, cl, sizes, alpha=c.alpha, beta=0.0)
    c._adjoints = (dA, dB)
    return c._adjoints


def _prec(prec) -> int:
    return PREC_BY_NAME[prec] if isinstance(prec, str) else int(prec)


class GettFunction:
    """Namespace holder; the Function class is created lazily so that importing
    tccg_amd does not import torch."""
    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            import torch

            class _Gett(torch.autograd.Function):
                @staticmethod
                def forward(ctx, A, B, contraction: Contraction, prec: int):
                    C = torch.empty(contraction.C0.torch_shape(), dtype=A.dtype, device=A.device)
                    contraction(A.contiguous(), B.contiguous(), C, alpha=contraction.alpha, beta=0.0, prec=prec)
                    ctx.save_for_backward(A, B)
                    ctx.contraction, ctx.prec = contraction, prec
                    return C

                @staticmethod
                def backward(ctx, dC):
                    A, B = ctx.saved_tensors
                    c, prec = ctx.contraction, ctx.prec
                    adjA, adjB = adjoints(c)
                    dC = dC.contiguous()
                    dA = dB = None
                    # gradients are written in the TCCG layout (contiguous),
                    # whatever the strides of the forward's inputs were
                    if ctx.needs_input_grad[0]:
                        dA = torch.empty(A.shape, dtype=A.dtype, device=A.device)
                        adjA(dC, B.contiguous(), dA, alpha=c.alpha, beta=0.0, prec=prec)
                    if ctx.needs_input_grad[1]:
                        dB = torch.empty(B.shape, dtype=B.dtype, device=B.device)
                        adjB(A.contiguous(), dC, dB, alpha=c.alpha, beta=0.0, prec=prec)
                    return dA, dB, None, None

            cls._fn = _Gett
        return cls._fn


def contract(A, B, contraction: Contraction, prec="f32", tune: bool = False):
    """alpha * contraction(A, B) as a differentiable torch op (C is a new tensor;
    the contraction's beta is not applied).  tune=True times the candidate
    plans of the forward and of both adjoint contractions on first use
    (Contraction.tune) and keeps the fastest."""
    p = _prec(prec)
    import torch
    want = torch.float64 if p == F64 else torch.float32
    if A.dtype != want or B.dtype != want:
        raise TypeError("precision %s expects %s inputs" % (prec, want))
    if tune and p not in contraction._tuned:
        for c in (contraction,) + adjoints(contraction):
            c.tune(p)
    return GettFunction.get().apply(A, B, contraction, p)
