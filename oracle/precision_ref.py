"""CPU emulations of the reduced-precision GETT forms (test infrastructure).

This file is part of the oracle: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it, as the checker.  The forms (split-BF16,
FP8 e4m3 with per-tensor scales) have no counterpart in the reference; each emulation
restates the arithmetic the GPU kernels are specified to perform, with torch
CPU casts for the roundings, and sums the products in FP64:

- split_bf16_parts / split_bf16_contract: x = x0 + x1 + x2 with
  x0 = bf16(x), x1 = bf16(x - x0), x2 = bf16(x - x0 - x1) (round to nearest
  even, each subtraction exact in FP32), and the six products
  a0b0 + a0b1 + a1b0 + a0b2 + a1b1 + a2b0.  Pinned by: the parts reconstruct
  FP32 values exactly (below the largest BF16); x0 equals torch's bf16 cast;
  hand-computed splits with ties; two-part inputs give the exact
  contraction, three-part inputs stay within the dropped products' bound,
  which a dropped or exchanged product breaks (tests/test_precision_ref.py).
- fp8_scale / fp8_quantize / fp8_contract: per-tensor scale s = 448 / max|x|
  (FP32 division), y = clamp(x * s, -448, 448) in FP32, OCP e4m3 by round to
  nearest even (torch.float8_e4m3fn), FP64 sum of the products, descaled by
  the FP32 value (max|A| / 448) * (max|B| / 448).  Pinned by: hand-computed
  e4m3 roundings including ties; exact round trips of e4m3 values; the
  contraction of exactly representable inputs equals einsum.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

from .einsum_ref import contract_ref


def _torch():
    import torch
    return torch


def split_bf16_parts(x: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The three round-to-nearest-even BF16 parts of FP32 `x`, as float64 arrays."""
    torch = _torch()
    t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
    p0 = t.to(torch.bfloat16).float()
    r1 = t - p0                       # exact in FP32
    p1 = r1.to(torch.bfloat16).float()
    r2 = r1 - p1                      # exact in FP32
    p2 = r2.to(torch.bfloat16).float()
    return p0.double().numpy(), p1.double().numpy(), p2.double().numpy()


# (part of the n-operand B, part of the m-operand A) of the six kept products
SPLIT_PRODUCTS = ((0, 0), (0, 1), (1, 0), (0, 2), (1, 1), (2, 0))


def split_bf16_contract(c: Sequence[str], a: Sequence[str], b: Sequence[str], A: np.ndarray, B: np.ndarray,
                        alpha: float = 1.0, grouped: bool = True) -> np.ndarray:
    """alpha * sum of the six kept part products, in FP64 (FP32 operands).
    grouped: the same six products as three contractions,
    a0 (b0 + b1 + b2) + a1 (b0 + b1) + a2 b0 (the part sums are exact in FP64),
    for full-size inputs; else one contraction per product."""
    pa, pb = split_bf16_parts(A), split_bf16_parts(B)
    if grouped:
        pairs = ((pa[0], pb[0] + pb[1] + pb[2]), (pa[1], pb[0] + pb[1]), (pa[2], pb[0]))
    else:
        pairs = tuple((pa[ia], pb[ib]) for ib, ia in SPLIT_PRODUCTS)
    out = None
    for xa, xb in pairs:
        term = contract_ref(c, a, b, xa, xb)
        out = term if out is None else out + term
    return alpha * out


E4M3_MAX = 448.0


def fp8_scale(x: np.ndarray) -> np.float32:
    """FP32 scale mapping max|x| onto 448 (1 for an all-zero tensor)."""
    amax = np.float32(np.abs(np.asarray(x, dtype=np.float32)).max()) if np.asarray(x).size else np.float32(0)
    if not amax > 0:
        amax = np.float32(1)
    return np.float32(np.float32(E4M3_MAX) / amax)


def fp8_quantize(x: np.ndarray, scale: np.float32) -> np.ndarray:
    """e4m3 values (as float64) of clamp(x * scale) computed in FP32."""
    torch = _torch()
    y = np.asarray(x, dtype=np.float32) * np.float32(scale)
    y = np.clip(y.astype(np.float32), np.float32(-E4M3_MAX), np.float32(E4M3_MAX))
    return torch.from_numpy(np.ascontiguousarray(y)).to(torch.float8_e4m3fn).double().numpy()


def fp8_descale(A: np.ndarray, B: np.ndarray) -> np.float32:
    """The epilogue's FP32 factor (max|A| / 448) * (max|B| / 448)."""
    def amax(x):
        m = np.float32(np.abs(np.asarray(x, dtype=np.float32)).max())
        return m if m > 0 else np.float32(1)
    return np.float32(np.float32(amax(A) / np.float32(E4M3_MAX)) * np.float32(amax(B) / np.float32(E4M3_MAX)))


def fp8_contract(c: Sequence[str], a: Sequence[str], b: Sequence[str], A: np.ndarray, B: np.ndarray,
                 alpha: float = 1.0) -> np.ndarray:
    """alpha * descale * sum of the e4m3 products, in FP64 (FP32 operands)."""
    qa = fp8_quantize(A, fp8_scale(A))
    qb = fp8_quantize(B, fp8_scale(B))
    return alpha * float(fp8_descale(A, B)) * contract_ref(c, a, b, qa, qb)
