"""LoG (loop over GEMM) on gfx950.

The reference's third contraction method (class GemmLoop,
tccg/gemm_loop.py:6-348).  The stride-one indices of A, B and C must take part
in one GEMM C[m, n] += A[m, k] B[k, n] over single indices: m is C's
stride-one index, n and k are B's and A's stride-one index when those are
free resp. contracted, or any free / contracted index otherwise
(tccg/gemm_loop.py:236-279).  Every other index is looped over: one free
index is the batch of a batched GEMM, the other free indices and the
contracted ones are explicit loops, and the contracted loops accumulate
(beta on their first iteration, 1 afterwards, tccg/gemm_loop.py:124-136).
Candidates are named as the reference names them (getCandidateName,
tccg/gemm_loop.py:61-76) and the ones with the largest GEMM, then the largest
batch, are kept (selectBestCandidate, tccg/gemm_loop.py:313-348).

On the GPU the instances of the GEMM are the problems of one grouped packed
launch (tccg_gett_group, tccg_amd.group): one problem per combination of the
free loops, all sharing one set of offset tables over (m, n, k) taken with
the full tensors' strides and differing only in their base pointers.  The
contracted loops are folded into the k range of every problem: its k table
runs over (contracted loop combination, k), so each C tile accumulates the
whole sum in registers and is written once (alpha and beta applied then),
where the reference accumulates into C from one GEMM call to the next.
Operands are packed once per distinct base: the instances that differ only
in an index an operand lacks share its packed planes.
"""
from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .plan import BF16, BF16X3, F32, F64, INVALID, PACKED_TILES

DUMMY = "dummy"


@dataclass
class LogCandidate:
    free_order: Tuple[str, ...]        # explicit free loops (outermost first)
    contracted_order: Tuple[str, ...]  # explicit contracted loops
    batch: str                         # free index of the batched GEMM, or DUMMY
    m: str
    n: str                             # DUMMY when the contraction has no n index
    k: str                             # DUMMY when it has no contracted index
    sizes: Tuple[int, int, int, int]   # batch, m, n, k extents (1 for DUMMY)

    @property
    def name(self) -> str:
        """getCandidateName (tccg/gemm_loop.py:61-76)."""
        f = "".join(self.free_order) if self.free_order else DUMMY
        c = "".join(self.contracted_order) if self.contracted_order else DUMMY
        return "LoG_free_%s_contracted_%s_batch_%s_%s%s%s" % (f, c, self.batch, self.m, self.n, self.k)

    def gemm_size(self) -> int:
        return self.sizes[1] * self.sizes[2] * self.sizes[3]


def log_candidates(pb) -> Optional[List[LogCandidate]]:
    """All LoG candidates of the (fused) problem, or None when LoG is not
    possible (a stride-one index that no single GEMM can hold,
    tccg/gemm_loop.py:254-258)."""
    AA, BB, C = pb.opA, pb.opB, pb.C            # AA holds C's stride-one index
    la, lb, lc = AA.labels(), BB.labels(), C.labels()
    size = {i.label: i.size for T in (AA, BB, C) for i in T.indices}
    m_set = [l for l in la if l in lc]
    n_set = [l for l in lb if l in lc]
    k_set = [l for l in la if l in lb]
    required = {la[0], lb[0], lc[0]}
    req_m = [l for l in m_set if l in required]
    req_n = [l for l in n_set if l in required]
    req_k = [l for l in k_set if l in required]
    if len(req_m) != 1 or len(req_n) >= 2 or len(req_k) >= 2:
        return None
    m = lc[0]
    n_choices = req_n if req_n else (n_set or [DUMMY])
    k_choices = req_k if req_k else (k_set or [DUMMY])
    out = []
    for n in n_choices:
        for k in k_choices:
            loop = [l for l in m_set + n_set + k_set if l not in (m, n, k)]
            free = [l for l in loop if l not in k_set]
            contracted = [l for l in loop if l in k_set]
            for batch in (free or [DUMMY]):
                rest = [l for l in free if l != batch]
                for fo in itertools.permutations(rest):
                    for co in itertools.permutations(contracted):
                        sz = lambda l: 1 if l == DUMMY else size[l]
                        out.append(LogCandidate(tuple(fo), tuple(co), batch, m, n, k,
                                                (sz(batch), sz(m), sz(n), sz(k))))
    return out


def best_log_candidates(cands: Sequence[LogCandidate], max_candidates: int) -> List[LogCandidate]:
    """Largest GEMM, then largest batch (tccg/gemm_loop.py:313-348)."""
    if not cands:
        return []
    g = max(c.gemm_size() for c in cands)
    b = max(c.sizes[0] for c in cands if c.gemm_size() == g)
    return [c for c in cands if c.gemm_size() == g and c.sizes[0] == b][:max(1, max_candidates)]


def _roundup(x: int, m: int) -> int:
    return -(-x // m) * m


def _lcm(a: int, b: int) -> int:
    import math
    return a * b // math.gcd(a, b)


def _table(values: np.ndarray, total: int) -> np.ndarray:
    out = np.full(total, INVALID, dtype=np.uint32)
    if np.any(values < 0) or np.any(values >= INVALID):
        raise ValueError("LoG element offsets beyond 32 bits")
    out[:values.size] = values.astype(np.uint32)
    return out


def log_tile(prec: int, Mg: int, Ng: int) -> int:
    """Packed tile with the least padded m x n area for the GEMM (the
    larger tile among equals)."""
    best, key = 0, None
    for t, (BM, BN, _) in enumerate(PACKED_TILES[prec]):
        area = _roundup(Mg, _lcm(BM, 64)) * _roundup(Ng, _lcm(BN, 64))
        k = (area, -BM * BN)
        if key is None or k < key:
            best, key = t, k
    return best


@dataclass
class LogLayout:
    """The grouped launch of one LoG candidate, host side: GEMM tables shared
    by every instance (k tables over (contracted combination, k)), the pack
    read orders, and the element offsets (opA, opB, C) of every instance."""
    prec: int
    tile: int
    tables: Tuple[np.ndarray, ...]            # offAm, offAk, offBn, offBk, offCm, offCn
    kfastA: int
    kfastB: int
    free_bases: List[Tuple[int, int, int]]

    @property
    def K(self) -> int:
        return int(np.count_nonzero(self.tables[1] != INVALID))


def log_layout(pb, cand: LogCandidate, prec: int = F32, tile: Optional[int] = None) -> LogLayout:
    AA, BB, CC = pb.opA, pb.opB, pb.C
    has = lambda T, l: l in T.labels()
    stride = lambda T, l: T.stride_of_label(l) if (l != DUMMY and has(T, l)) else 0
    ext = lambda l: 1 if l == DUMMY else next(i.size for X in (AA, BB, CC) for i in X.indices if i.label == l)
    Mg, Ng, Kg = ext(cand.m), ext(cand.n), ext(cand.k)
    if tile is None:
        tile = log_tile(prec, Mg, Ng)
    BM, BN, BK = PACKED_TILES[prec][tile]

    def bases(labels):
        combos = [(0, 0, 0)]
        for l in labels:
            sa, sb, sc = stride(AA, l), stride(BB, l), stride(CC, l)
            combos = [(x + i * sa, y + i * sb, z + i * sc) for i in range(ext(l)) for (x, y, z) in combos]
        return combos

    cb = np.array(bases(list(cand.contracted_order)), dtype=np.int64).reshape(-1, 3)
    kk = np.arange(Kg, dtype=np.int64)
    Kp = _roundup(cb.shape[0] * Kg, _lcm(BK, 32))
    Mp, Np = _roundup(Mg, _lcm(BM, 64)), _roundup(Ng, _lcm(BN, 64))
    mm, nn = np.arange(Mg, dtype=np.int64), np.arange(Ng, dtype=np.int64)
    offAk = _table((cb[:, 0:1] + kk[None, :] * stride(AA, cand.k)).reshape(-1), Kp)
    offBk = _table((cb[:, 1:2] + kk[None, :] * stride(BB, cand.k)).reshape(-1), Kp)
    offAm, offCm = _table(mm * stride(AA, cand.m), Mp), _table(mm * stride(CC, cand.m), Mp)
    offBn, offCn = _table(nn * stride(BB, cand.n), Np), _table(nn * stride(CC, cand.n), Np)
    # gather along each operand's stride-one index (the pack kernel's read order)
    kfastA = 1 if (cand.k != DUMMY and AA.labels()[0] == cand.k) else 0
    kfastB = 1 if (cand.k != DUMMY and BB.labels()[0] == cand.k) else 0
    free = ([cand.batch] if cand.batch != DUMMY else []) + list(cand.free_order)
    return LogLayout(prec, tile, (offAm, offAk, offBn, offBk, offCm, offCn), kfastA, kfastB, bases(free))


_TORCH_DTYPE = {F32: "float32", F64: "float64", BF16X3: "float32", BF16: "float32"}


class DeviceLoG:
    """One LoG candidate bound to fixed tensors: every GEMM instance a problem
    of one grouped packed launch (a pack launch and a GETT launch)."""

    def __init__(self, contraction, cand: LogCandidate, A, B, C, alpha: Optional[float] = None,
                 beta: Optional[float] = None, prec: int = F32, tile: Optional[int] = None, order: str = "xcd"):
        import torch
        from .group import GettGroup
        if prec not in _TORCH_DTYPE:
            raise ValueError("LoG runs FP32, FP64, split-BF16 and BF16")
        pb = contraction.problem
        self.contraction, self.cand, self.prec = contraction, cand, prec
        self.alpha = contraction.alpha if alpha is None else alpha
        self.beta = contraction.beta if beta is None else beta
        dt = getattr(torch, _TORCH_DTYPE[prec])
        for name, t, T in (("A", A, contraction.A0), ("B", B, contraction.B0), ("C", C, contraction.C0)):
            if (t.dtype != dt or not t.is_contiguous() or t.numel() != T.size() or t.device.type != "cuda"
                    or t.device != C.device):
                raise ValueError("LoG: %s must be a contiguous %s tensor of %d elements on C's GPU"
                                 % (name, _TORCH_DTYPE[prec], T.size()))
        opA, opB = (B, A) if pb.swapped else (A, B)
        lay = log_layout(pb, cand, prec, tile)
        self.layout = lay
        offAm, offAk, offBn, offBk, offCm, offCn = lay.tables
        esz = A.element_size()
        g = GettGroup(prec, lay.tile, order)
        for fa, fb, fc in lay.free_bases:
            a = g.operand(opA.data_ptr() + esz * fa, offAm, offAk, lay.kfastA, False, key=("A", fa))
            b = g.operand(opB.data_ptr() + esz * fb, offBn, offBk, lay.kfastB, True, key=("B", fb))
            g.problem(a, b, C.data_ptr() + esz * fc, offCm, offCn, self.alpha, self.beta, locality=fa)
        self.group = g.bind(C.device)
        self.nlaunches = 2
        self._tensors = (A, B, C)

    def launch(self, stream=None) -> None:
        self.group.launch(stream)

    def run(self, A=None, B=None, C=None, alpha=None, beta=None, stream=None) -> None:
        """Launch on the bound tensors with the bound alpha and beta (the
        arguments exist for the timing helpers' calling convention)."""
        for t, b in zip((A, B, C), self._tensors):
            if t is not None and t.data_ptr() != b.data_ptr():
                raise ValueError("LoG runs on the tensors it was bound to")
        if (alpha is not None and alpha != self.alpha) or (beta is not None and beta != self.beta):
            raise ValueError("LoG runs with the alpha and beta it was bound to")
        self.launch(stream)
