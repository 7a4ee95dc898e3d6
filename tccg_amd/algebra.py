"""Index classification of the GETT planner.

Restates the reference's m/n/k/loop classification
(tccg/tccg_util.py:212-296, tccg/tccg_util.py:597-611).  (The index-set
splitting and block-size candidates of the reference's AVX2 planner,
tccg/tccg_util.py:307-552 and tccg/gett.py:689-698, live with the AVX2
restatement in oracle/avx2_model.py; the gfx950 planner blocks index sets
through its offset tables instead, tccg_amd/plan.py.)
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

from .ir import Index, Tensor


# ---------------------------------------------------------------------------
# m / n / k classification
# ---------------------------------------------------------------------------
def get_loop_indices(A: Tensor, B: Tensor, C: Tensor) -> List[Index]:
    """Indices present in A, B and C (batch indices), in B's order
    (tccg/tccg_util.py:212-227)."""
    return [i for i in B.indices if C.has_index(i) and A.has_index(i)]


def _holder_and_other(A: Tensor, B: Tensor, C: Tensor) -> Tuple[Tensor, Tensor]:
    """The input that holds C's stride-one index first (tccg/tccg_util.py:233-238)."""
    if A.has_index(C.indices[0]):
        return A, B
    return B, A


def get_m_indices(A: Tensor, B: Tensor, C: Tensor) -> List[Index]:
    """C's indices (in C order) that are in the input holding C's stride-one index
    and not in the other input (tccg/tccg_util.py:276-296)."""
    X, Y = _holder_and_other(A, B, C)
    return [i for i in C.indices if X.has_index(i) and not Y.has_index(i)]


def get_n_indices(A: Tensor, B: Tensor, C: Tensor) -> List[Index]:
    """C's indices (in C order) that are only in the other input
    (tccg/tccg_util.py:230-250)."""
    X, Y = _holder_and_other(A, B, C)
    return [i for i in C.indices if Y.has_index(i) and not X.has_index(i)]


def get_contracted_indices(A: Tensor, B: Tensor, C: Tensor) -> List[Index]:
    """A's indices (in A order) present in B and absent from C
    (tccg/tccg_util.py:597-611)."""
    return [i for i in A.indices if B.has_index(i) and not C.has_index(i)]


def product(indices: Sequence[Index]) -> int:
    n = 1
    for i in indices:
        n *= i.size
    return n
