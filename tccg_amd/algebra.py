"""Index-set algebra of the GETT planner.

Restates the reference's index classification and splitting
(tccg/tccg_util.py:179-334, tccg/tccg_util.py:382-552,
tccg/tccg_util.py:597-611) and the block-size candidates of the GETT
generator (tccg/gett.py:689-698).
"""
from __future__ import annotations

import copy
import itertools
from typing import Dict, List, Sequence, Tuple

from .ir import Index, Tensor, has_item


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


# ---------------------------------------------------------------------------
# prime factors and index splitting
# ---------------------------------------------------------------------------
def _primes_below(limit: int) -> List[int]:
    sieve = bytearray([1]) * (limit + 1)
    sieve[0:2] = b"\x00\x00"
    for p in range(2, int(limit ** 0.5) + 1):
        if sieve[p]:
            sieve[p * p :: p] = bytearray(len(sieve[p * p :: p]))
    return [p for p in range(limit + 1) if sieve[p]]


_PRIMES_5000 = _primes_below(5000)


def prime_factors(a: int) -> List[int]:
    """Prime factors of a that are at most 5000, with multiplicity, ascending;
    [1] for a == 1 (tccg/tccg_util.py:316-334: trial division stops at 5000, so
    a larger prime factor is not reported)."""
    if a == 1:
        return [1]
    out = []
    for p in _PRIMES_5000:
        if p > a:
            break
        while a != 1 and a % p == 0:
            out.append(p)
            a //= p
    return out


def split_index(idx: Index, size: int) -> Tuple[Index, Index]:
    """(label_0: size, label_1: extent/size) (tccg/tccg_util.py:307-314)."""
    if idx.size % size != 0:
        raise ValueError("Index %s: nothing to split." % str(idx))
    return Index(idx.label + "_0", size), Index(idx.label + "_1", idx.size // size)


def _find(item: Index, items: Sequence[Index]) -> int:
    for i, x in enumerate(items):
        if x == item:
            return i
    return -1


def _merge_same_label(items: Sequence[Index]) -> List[Index]:
    out: List[Index] = []
    for idx in items:
        p = _find(idx, out)
        if p >= 0:
            out[p].size *= idx.size
        else:
            out.append(idx)
    return out


def _candidate_exists(sol: Sequence[Index], candidates: Sequence[Sequence[Index]]) -> bool:
    for old in candidates:
        same = all(any(o.size == n.size and o.label == n.label for o in old) for n in sol)
        if same and len(old) == len(sol):
            return True
    return False


def _solutions(primes: List[int], counts: Dict[int, int], with_factor: Dict[int, List[Index]],
               solutions: List[List[Index]]) -> List[List[Index]]:
    """All ways to draw the prime factors of the target size from the indices'
    own factorisations (tccg/tccg_util.py:382-444)."""
    for p in primes:
        unique: List[List[Index]] = []
        for subset in itertools.combinations(with_factor[p], counts[p]):
            occ: Dict[str, int] = {}
            for idx in subset:
                occ[idx.label] = occ.get(idx.label, 0) + 1
            sol: List[Index] = []
            for idx in subset:
                cand = Index(idx.label, p ** occ[idx.label])
                if not has_item(sol, cand):
                    sol.append(cand)
            if not _candidate_exists(sol, unique):
                unique.append(sol)
        combined: List[List[Index]] = []
        for s1 in solutions:
            for s2 in unique:
                s = _merge_same_label(copy.deepcopy(s1 + s2))
                if not _candidate_exists(s, combined):
                    combined.append(s)
        solutions = combined
    return solutions


SplitSolution = Tuple[List[Index], List[Index], Tensor, Tensor]


def split_index_set(indices: Sequence[Index], size: int, tensorA: Tensor, keep_stride1_A: bool,
                    tensorB: Tensor, keep_stride1_B: bool, register_size: int = 8,
                    B_is_important: bool = False) -> List[SplitSolution]:
    """Split an index set I into I0 (extent product == size) and I1 = I \\ I0,
    splitting individual indices where needed (tccg/tccg_util.py:463-552).

    A candidate is valid when, for each of tensorA/tensorB whose stride-one
    index belongs to I, that index is kept in I0 (if requested), and, when
    B_is_important, tensorB's stride-one part in I0 is a multiple of
    register_size.  Returns [(I0, I1, tensorA', tensorB')] with split indices
    replaced in both tensors; [] when no valid split exists.
    """
    indices = list(indices)
    primes = prime_factors(size)
    counts: Dict[int, int] = {}
    with_factor: Dict[int, List[Index]] = {}
    for p in primes:
        counts[p] = counts.get(p, 0) + 1
        with_factor[p] = []
    for idx in indices:
        for p in prime_factors(idx.size):
            if p in with_factor:
                with_factor[p].append(idx)
    candidates = _solutions(sorted(set(primes)), counts, with_factor, [[]])

    valid = []
    a0, b0 = tensorA.indices[0], tensorB.indices[0]
    for cand in candidates:
        okA = (not has_item(indices, a0)) or (not keep_stride1_A) or has_item(cand, a0)
        okB = (not has_item(indices, b0)) or (not keep_stride1_B) or has_item(cand, b0)
        if not (okA and okB):
            continue
        pb = _find(b0, cand)
        if B_is_important and pb != -1 and cand[pb].size % register_size != 0:
            continue
        valid.append(cand)

    out: List[SplitSolution] = []
    for cand in valid:
        ind0: List[Index] = []
        ind1: List[Index] = []
        tA = copy.deepcopy(tensorA)
        tB = copy.deepcopy(tensorB)
        for idx in cand:
            p = _find(idx, indices)
            if p == -1:
                raise ValueError("split_index_set(): index %s not found" % idx)
            whole = indices[p]
            if idx.size != whole.size:
                i0, i1 = split_index(whole, idx.size)
                if not (tA.has_index(whole) and tB.has_index(whole)):
                    raise ValueError("split_index_set(): index %s not found in tensor" % whole)
                tA.replace_index(whole, [i0, i1])
                tB.replace_index(whole, [i0, i1])
                ind0.append(i0)
                ind1.append(i1)
            else:
                ind0.append(idx)
        for idx in indices:
            if not has_item(cand, idx):
                ind1.append(idx)
        out.append((ind0, ind1, tA, tB))
    return out


def get_possible_values(block: int, must_divide: int, target: int, max_num: int) -> List[int]:
    """Up to max_num multiples of `block` that divide `must_divide`, descending
    from about `target` (tccg/gett.py:689-698, integer division as in the
    reference)."""
    out: List[int] = []
    a = target // block
    while len(out) < max_num and a >= 1:
        while must_divide % (a * block) != 0 and a > 1:
            a -= 1
        if a * block > 0 and must_divide % (a * block) == 0:
            out.append(a * block)
        a -= 1
    return out


# ---------------------------------------------------------------------------
# permutations between a sub-tensor and its packed form
# ---------------------------------------------------------------------------
def get_perm(IN: Tensor, OUT: Tensor) -> List[int]:
    """perm[j] = position in IN of OUT's j-th index (tccg/tccg_util.py:186-191)."""
    return [IN.pos(i) for i in OUT.indices]


def outer_sizes(ld: Sequence[int], last_extent: int) -> List[int]:
    """Ratios of consecutive strides plus the last extent: the padded extents of a
    strided sub-tensor (tccg/tccg_util.py:179-184)."""
    return [ld[i] // ld[i - 1] for i in range(1, len(ld))] + [last_extent]


def transpose_args(IN: Tensor, OUT: Tensor):
    """(perm, size, outer sizes of IN, outer sizes of OUT) of the packing
    permutation IN -> OUT (tccg/tccg_util.py:193-203)."""
    perm = get_perm(IN, OUT)
    size = [i.size for i in IN.indices]
    return perm, size, outer_sizes(IN.ld, IN.indices[-1].size), outer_sizes(OUT.ld, OUT.indices[-1].size)
