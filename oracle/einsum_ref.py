"""numpy.einsum FP64 oracle on TCCG-layout arrays (test infrastructure only).

A TCCG tensor X[i0, ..., ir] (leftmost index stride-one, tccg/tensor.py:8-13)
is held by a C-contiguous array of shape (ir, ..., i0); the einsum subscripts
are therefore the reversed index lists.  Computes
    C = alpha * sum_k A * B + beta * C
in float64, the semantics of the reference's loop nest (tccg/tccg.py:686-723).
"""
from __future__ import annotations

import string
from typing import Dict, Sequence

import numpy as np


def subscripts(c: Sequence[str], a: Sequence[str], b: Sequence[str]) -> str:
    letters: Dict[str, str] = {}
    pool = iter(string.ascii_letters)
    for lab in list(c) + list(a) + list(b):
        if lab not in letters:
            letters[lab] = next(pool)
    rev = lambda idx: "".join(letters[l] for l in reversed(list(idx)))
    return "%s,%s->%s" % (rev(a), rev(b), rev(c))


def contract_ref(c, a, b, A, B, C=None, alpha: float = 1.0, beta: float = 0.0) -> np.ndarray:
    """FP64 result in TCCG layout (array shape reversed(c))."""
    A64 = np.asarray(A, dtype=np.float64)
    B64 = np.asarray(B, dtype=np.float64)
    out = np.einsum(subscripts(c, a, b), A64, B64, optimize=True)
    if alpha != 1.0:
        out = alpha * out
    if beta != 0.0:
        if C is None:
            raise ValueError("beta != 0 needs C")
        out = out + beta * np.asarray(C, dtype=np.float64)
    return out


def transpose_ref(in_labels: Sequence[str], out_labels: Sequence[str], X, alpha: float = 1.0, beta: float = 0.0,
                  OUT=None) -> np.ndarray:
    """out[out_labels] = alpha * X[in_labels] + beta * OUT in float64, TCCG
    layout on both sides: the semantics of the reference's Transpose block
    (tccg/transpose.py:74-144: an HPTT plan with perm[j] = position in the
    input of output index j, getPerm, tccg/tccg_util.py:186-191, applied as
    B = alpha * permute(A) + beta * B)."""
    in_labels, out_labels = list(in_labels), list(out_labels)
    rev_in = list(reversed(in_labels))
    axes = [rev_in.index(l) for l in reversed(out_labels)]
    out = alpha * np.transpose(np.asarray(X, dtype=np.float64), axes)
    if beta != 0.0:
        if OUT is None:
            raise ValueError("beta != 0 needs OUT")
        out = out + beta * np.asarray(OUT, dtype=np.float64)
    return np.ascontiguousarray(out)


def rel_frobenius(x, ref) -> float:
    """||x - ref||_F / ||ref||_F, in float64."""
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.linalg.norm(ref.ravel())
    num = np.linalg.norm((x - ref).ravel())
    return float(num / den) if den > 0 else float(num)


def reference_equal(A, B, is_double: bool) -> bool:
    """The reference harness's element check (tccg/tccg.py:843-881):
    per element min(|d|/max(|a|,|b|), |d|) must not exceed 1e-4 (float) or
    1e-9 (double).  The reference compares |a| with |b| (its `diff` is the
    difference of absolute values); this restatement keeps that."""
    a = np.abs(np.asarray(A, dtype=np.float64).ravel())
    b = np.abs(np.asarray(B, dtype=np.float64).ravel())
    mx = np.maximum(a, b)
    diff = np.abs(a - b)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(diff > 0, np.minimum(diff / mx, diff), 0.0)
    return bool(np.all(rel <= (1e-9 if is_double else 1e-4)))


def reference_inputs(n: int, which: str, dtype=np.float32) -> np.ndarray:
    """Deterministic inputs of the reference harness (tccg/tccg.py:957-971)."""
    i = np.arange(n, dtype=np.int64)
    if which == "A":
        v = (10 + ((i + 1) * 17) % 100) / 110.0
    elif which == "B":
        v = (10 + ((i + 2) * 3) % 100) / 110.0
    elif which == "C":
        v = (10 + ((i + 3) * 13) % 100) / 110.0
    else:
        raise ValueError(which)
    # the reference computes (float)(...) / 110.0 in double, then stores as T
    return v.astype(np.float32).astype(dtype) if dtype == np.float32 else v.astype(dtype)
