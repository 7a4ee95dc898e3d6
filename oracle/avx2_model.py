"""The reference's AVX2 GETT planner: index-set splitting, block-size
candidates, the performance model and the model-ranked candidate list
(test infrastructure and CPU baseline only).

Restates, in code of its own:
- the prime factorisation and index-set splitting (tccg/tccg_util.py:307-334,
  tccg/tccg_util.py:382-552) and getPerm (tccg/tccg_util.py:186-191);
- getPossibleValues (tccg/gett.py:689-698);
- the avx2 architecture constants (tccg/arch.py:21-55);
- Gett.estimateGFLOPS (tccg/gett.py:590-687), with Python 2 integer
  division where the reference divides integers;
- the candidate search of Gett.genCode (tccg/gett.py:961-1150): both loop
  variants (tccg/gett.py:52-55), mc/nc/kc from getPossibleValues, every
  split of the m, n and k sets for L2 and L1, the best-rated permutation per
  (variant, mc, nc, kc, mc1, nc1, mr, nr) key, only the largest macro-kernel
  among keys of one variant with equal estimates, then the top
  maxImplementations.  Among equal estimates the reference's order follows
  Python 2 dict iteration, which no file pins; here ties sort by variant,
  then larger mc*nc*kc, then (mc, nc, kc).
"""
from __future__ import annotations

import copy
import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from tccg_amd.ir import Index, Tensor, has_item


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


# ---------------------------------------------------------------------------
# the avx2 architecture (tccg/arch.py:21-55) and the GETT loop variants
# ---------------------------------------------------------------------------
class Avx2Arch:
    """One Haswell core as the reference models it."""

    def __init__(self, float_type: str = "float"):
        self.float_size = 8 if float_type == "double" else 4
        self.axpy_bandwidth = 13.0 * 2 ** 30           # B/s
        self.register_size = 256 // 8 // self.float_size  # elements per __m256
        self.cache_line_size = 64 // self.float_size      # elements
        self.frequency = 2.5
        self.num_fma_cycle = 2
        self.L1_SIZE = 32 * 1024.0
        self.L2_SIZE = 256 * 1024.0
        self.L1_ASSOCIATIVITY = 8
        self.L2_ASSOCIATIVITY = 8

    def peak_gflops(self) -> float:
        return self.frequency * self.num_fma_cycle * self.register_size * 2


# loop order and packing positions, leftmost outermost (tccg/gett.py:52-55)
VARIANTS = (
    ["n", "k", "B", "m", "A", "kernel_NC_MC", "C", "}m", "}k", "}n"],   # 0: panel-matrix (n outer)
    ["m", "k", "A", "n", "B", "kernel_MC_NC", "C", "}n", "}k", "}m"],   # 1: matrix-panel (m outer)
)


class Avx2Gett:
    """The reference GETT generator's view of one contraction (tccg/gett.py:18-89):
    m/n/k sets, the A/B exchange that makes A hold the m set, M, N, K."""

    def __init__(self, A: Tensor, B: Tensor, C: Tensor, float_type: str = "float", beta: float = 0.0):
        from tccg_amd.algebra import get_contracted_indices, get_m_indices, get_n_indices
        self.arch = Avx2Arch(float_type)
        self.float_type = float_type
        self.float_size = self.arch.float_size
        self.beta = beta
        self.mInd = get_m_indices(A, B, C)
        self.nInd = get_n_indices(A, B, C)
        self.kInd = get_contracted_indices(A, B, C)
        self.C = copy.deepcopy(C)
        if A.has_index(self.nInd[0]):
            self.A, self.B = copy.deepcopy(B), copy.deepcopy(A)
        else:
            self.A, self.B = copy.deepcopy(A), copy.deepcopy(B)
        prod = lambda xs: int(__import__("numpy").prod([i.size for i in xs])) if xs else 1
        self.sizeM, self.sizeN, self.sizeK = prod(self.mInd), prod(self.nInd), prod(self.kInd)

    def micro_tile(self) -> Tuple[int, int]:
        """(mr, nr) of the avx2 micro-kernel (tccg/gett.py:1008-1015)."""
        return (12, 4) if self.float_type == "double" else (24, 4)

    def estimate_gflops(self, mc: int, nc: int, kc: int, mc1: int, nc1: int, mr: int, nr: int,
                        permA: Sequence[int], Ahat: Tensor, permB: Sequence[int], Bhat: Tensor, Chat: Tensor,
                        variant: Sequence[str]) -> float:
        """Gett.estimateGFLOPS (tccg/gett.py:590-687): peak * T_gemm / (T_pack + T_gemm)."""
        arch = self.arch
        num_iter = {"m": (self.sizeM + mc - 1) // mc, "n": (self.sizeN + nc - 1) // nc,
                    "k": (self.sizeK + kc - 1) // kc}
        size = {"A": mc * kc, "B": nc * kc, "C": mc * nc}
        perm = {"A": list(permA), "B": list(permB), "C": [0]}
        contiguous = {"A": Ahat.count_contiguous_stride_one_elements(),
                      "B": Bhat.count_contiguous_stride_one_elements(),
                      "C": Chat.count_contiguous_stride_one_elements()}
        vec_eff = {"A": 1.0, "B": 1.0, "C": 1.0}
        rs = arch.register_size
        if Ahat.indices[0].size < rs or Ahat.indices[permA[0]].size < rs:
            vec_eff["A"] = 0.95
        if Bhat.indices[0].size < rs or Bhat.indices[permB[0]].size < rs:
            vec_eff["B"] = 0.95
        packing = 0.0
        for tensor in ("A", "B", "C"):
            open_loops: List[str] = []
            for token in variant:
                if token == tensor:
                    break
                if token.startswith("}"):
                    open_loops.remove(token[1])
                elif token in ("m", "n", "k"):
                    open_loops.append(token)
            moved = size[tensor] * float(self.float_size)
            for it in open_loops:
                moved *= num_iter[it]
            # (len(perm) - 1) counts the three tensors of the dict: a constant 0.97
            bandwidth = arch.axpy_bandwidth * (1 - (len(perm) - 1) * 0.015)
            n = contiguous[tensor]
            cl_eff = n / (float((n + arch.cache_line_size - 1) // arch.cache_line_size) * arch.cache_line_size)
            t = moved / (bandwidth * cl_eff * vec_eff[tensor])
            if perm[tensor][0] != 0:
                t = moved / (bandwidth * 0.8 * cl_eff * vec_eff[tensor])
            packing += (3 if (tensor == "C" and self.beta != 0) else 2) * t
        eff = 0.95
        l2 = ((arch.L2_ASSOCIATIVITY - 1.0) / arch.L2_ASSOCIATIVITY) * arch.L2_SIZE
        if variant[0] == "m":
            if (nc * kc + 2 * mc1 * kc) * self.float_size > l2:
                eff *= 0.85
        else:
            if (mc * kc + 2 * nc1 * kc) * self.float_size > l2:
                eff *= 0.85
        outer = next(tok.split("_")[1] for tok in variant if tok.startswith("kernel"))
        copiesA, copiesB = (2, 1) if outer == "MC" else (1, 2)
        req_l1 = (copiesB * nc1 * kc + copiesA * mc1 * kc) * self.float_size
        if req_l1 > ((arch.L1_ASSOCIATIVITY - 1.0) / arch.L1_ASSOCIATIVITY) * arch.L1_SIZE:
            eff *= 0.85
        flops = self.sizeM * self.sizeN * self.sizeK * 2
        t_gemm = flops / 1e9 / (arch.peak_gflops() * eff)
        return arch.peak_gflops() * t_gemm / (packing + t_gemm)


@dataclass
class Avx2Candidate:
    variant: int
    mc: int
    nc: int
    kc: int
    mc1: int
    nc1: int
    mr: int
    nr: int
    indices: str
    estimate: float

    def key(self) -> Tuple[int, ...]:
        return (self.variant, self.mc, self.nc, self.kc, self.mc1, self.nc1, self.mr, self.nr)

    def name(self) -> str:
        """tccg/gett.py:578-579."""
        return "gett_var%d_mc%d_nc%d_kc%d_mc1%d_nc1%d_mr%d_nr%d_%s" % (self.key() + (self.indices,))


def enumerate_candidates(g: Avx2Gett) -> List[Avx2Candidate]:
    """Every (variant, mc, nc, kc, split) the reference's genCode estimates
    (tccg/gett.py:1003-1110), with its estimated GFLOP/s."""
    mr, nr = g.micro_tile()
    mc1, nc1 = mr, nr
    dbl = g.float_type == "double"
    out: List[Avx2Candidate] = []
    rs = g.arch.register_size
    for vid, variant in enumerate(VARIANTS):
        mc_values = get_possible_values(mr, g.sizeM, 96 if dbl else 192, 3)
        if variant[0] == "m":
            mc_values = [g.sizeM] if g.sizeM < 4600 else get_possible_values(mr, g.sizeM, 4600, 3)
        for mc in mc_values:
            if g.sizeM % mc or mc % mr:
                continue
            if g.sizeN % nc1 or nc1 % nr:
                continue
            nc_values = get_possible_values(nc1, g.sizeN, 96 if dbl else 192, 3)
            if variant[0] == "n":
                nc_values = [g.sizeN] if g.sizeN < 4600 else get_possible_values(nc1, g.sizeN, 4600, 3)
            for nc in nc_values:
                if g.sizeN % nc:
                    continue
                for kc in get_possible_values(1, g.sizeK, 256, 4):
                    if g.sizeK % kc:
                        continue
                    for mInd0, mInd1, tA, tC in split_index_set(copy.deepcopy(g.mInd), mc, g.A, True, g.C, True,
                                                                rs, True):
                        for nInd0, nInd1, tB, tC2 in split_index_set(copy.deepcopy(g.nInd), nc, g.B, True, tC,
                                                                     True):
                            for kInd0, kInd1, tA2, tB2 in split_index_set(copy.deepcopy(g.kInd), kc, tA, True, tB,
                                                                          True):
                                for mL1, mRem, tA3, tC3 in split_index_set(copy.deepcopy(mInd0), mc1, tA2, False,
                                                                           tC2, True, rs, True):
                                    for nL1, nRem, tB3, tC4 in split_index_set(copy.deepcopy(nInd0), nc1, tB2,
                                                                               False, tC3, True):
                                        Ahat = tA3.sub_tensor(mL1 + mRem + kInd0)
                                        Bhat = tB3.sub_tensor(nL1 + nRem + kInd0)
                                        Chat = tC4.sub_tensor(mL1 + mRem + nL1 + nRem)
                                        Atilde = Tensor("A~", mL1 + kInd0 + mRem)
                                        Btilde = Tensor("B~", nL1 + kInd0 + nRem)
                                        permA = get_perm(Ahat, Atilde)
                                        permB = get_perm(Bhat, Btilde)
                                        est = g.estimate_gflops(mc, nc, kc, mc1, nc1, mr, nr, permA, Ahat, permB,
                                                                Bhat, Chat, variant)
                                        ind = "%s_%s_%s" % (tA3.index_str(), tB3.index_str(), tC4.index_str())
                                        out.append(Avx2Candidate(vid, mc, nc, kc, mc1, nc1, mr, nr, ind, est))
    return out


def rank_candidates(cands: Sequence[Avx2Candidate], max_implementations: int = 16) -> List[Avx2Candidate]:
    """The candidates genCode generates (tccg/gett.py:985-1000, 1113-1150): the
    best-rated split per key, only the largest mc*nc*kc among keys of one
    variant with equal estimates, the top max_implementations (-1: all)."""
    best: Dict[Tuple[int, ...], Avx2Candidate] = {}
    for c in cands:
        if c.key() not in best or c.estimate > best[c.key()].estimate:
            best[c.key()] = c
    kept = []
    for k, c in best.items():
        largest = max(o.mc * o.nc * o.kc for o in best.values()
                      if o.variant == c.variant and o.estimate == c.estimate)
        if c.mc * c.nc * c.kc >= largest:
            kept.append(c)
    kept.sort(key=lambda c: (-c.estimate, c.variant, -(c.mc * c.nc * c.kc), c.mc, c.nc, c.kc))
    return kept if max_implementations < 0 else kept[:max_implementations]


def candidates_for(c_labels: Sequence[str], a_labels: Sequence[str], b_labels: Sequence[str],
                   sizes: Dict[str, int], float_type: str = "float", beta: float = 0.0,
                   max_implementations: int = 16) -> List[Avx2Candidate]:
    mk = lambda lab, idx: Tensor(lab, [Index(l, sizes[l]) for l in idx])
    g = Avx2Gett(mk("A", a_labels), mk("B", b_labels), mk("C", c_labels), float_type, beta)
    return rank_candidates(enumerate_candidates(g), max_implementations)
