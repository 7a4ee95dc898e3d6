"""TTGT on gfx950: transpose the operands into matrix layouts, multiply, and
transpose the product back.

The reference's second contraction method (class TTGEMMT,
tccg/ttgemmt.py:133-336).  For every order of the m, n and k index sets,
every A/B exchange and every operand orientation (N or T) it writes a
candidate made of up to three transpositions around one GEMM:

  A~ = transpose(A) to (m, k) or (k, m)          unless A already is laid out so
  B~ = transpose(B) to (k, n) or (n, k)          unless B already is
  C~ = alpha * A~ B~                              (C itself, with beta, when C is (m, n))
  C  = transpose(C~) + beta * C                  when C is not laid out as (m, n)

scores each by the transposition time of its performance model
(Transpose.getTransposeTime, tccg/transpose.py:96-101: bytes over the
copy bandwidth, 0.71 of it when the stride-one index moves) and keeps those
of minimal cost (selectBestCandidate, tccg/ttgemmt.py:96-114).  Here the
transpositions are the HIP kernels of tccg_amd.transpose and the GEMM is
this backend's own GETT over the transposed temporaries (after index fusion
a plain matrix product).  Temporaries live in one workspace, as the
reference's allocateMemory (tccg/tccg_util.py:149-165) places them.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

from .plan import F32, F64

_ALIGN = 256


@dataclass
class TtgtCandidate:
    m: Tuple[str, ...]          # m index order of the GEMM (after the exchange)
    n: Tuple[str, ...]
    k: Tuple[str, ...]
    swap: int                   # 1: the GEMM's left operand is the contraction's B
    opA: str                    # "N": left operand laid out (m, k); "T": (k, m)
    opB: str                    # "N": right operand (k, n); "T": (n, k)
    idxA: List[str]             # index order of the left GEMM operand
    idxB: List[str]
    idxC: List[str]             # (m, n): the GEMM's output
    transA: bool
    transB: bool
    transC: bool
    cost: float                 # model transposition time (seconds)

    @property
    def name(self) -> str:
        """The reference's candidate name (getCandidateName, tccg/ttgemmt.py:116-127)."""
        return "%s_%s_%s_%d%s%s" % ("".join(self.m), "".join(self.n), "".join(self.k), self.swap, self.opA, self.opB)


def _transpose_time(in_labels: Sequence[str], out_labels: Sequence[str], elements: int, esz: int,
                    bandwidth: float) -> float:
    """Transpose.getTransposeTime (tccg/transpose.py:92-101): the input's
    bytes over the copy bandwidth, over 0.71 of it when the output's first
    index is not the input's first."""
    moved = esz * elements
    return moved / (bandwidth * (0.71 if out_labels[0] != in_labels[0] else 1.0))


def ttgt_candidates(pb, esz: int = 4, bandwidth: float = 8.0e12) -> List[TtgtCandidate]:
    """Every TTGT candidate of the (fused) problem with its model cost, in the
    reference's enumeration order (tccg/ttgemmt.py:168-203)."""
    A, B, C = pb.opA, pb.opB, pb.C
    la, lb, lc = A.labels(), B.labels(), C.labels()
    sizes = {i.label: i.size for T in (A, B, C) for i in T.indices}
    m = [i.label for i in pb.m]
    n = [i.label for i in pb.n]
    k = [i.label for i in pb.k]
    nel = lambda labels: _prod(sizes[l] for l in labels)
    out = []
    for mi in itertools.permutations(m):
        for ni in itertools.permutations(n):
            for ki in itertools.permutations(k):
                for swap in (0, 1):
                    X, Y = (B, A) if swap else (A, B)
                    lx, ly = (lb, la) if swap else (la, lb)
                    m_, n_ = (ni, mi) if swap else (mi, ni)
                    idxC = list(m_) + list(n_)
                    for opA in ("N", "T"):
                        idxA = list(m_) + list(ki) if opA == "N" else list(ki) + list(m_)
                        for opB in ("N", "T"):
                            idxB = list(ki) + list(n_) if opB == "N" else list(n_) + list(ki)
                            tA, tB, tC = lx != idxA, ly != idxB, lc != idxC
                            cost = 0.0
                            if tA:
                                cost += _transpose_time(lx, idxA, nel(lx), esz, bandwidth)
                            if tB:
                                cost += _transpose_time(ly, idxB, nel(ly), esz, bandwidth)
                            if tC:
                                cost += _transpose_time(idxC, lc, nel(lc), esz, bandwidth)
                            out.append(TtgtCandidate(tuple(m_), tuple(n_), tuple(ki), swap, opA, opB, idxA, idxB,
                                                     idxC, tA, tB, tC, cost))
    return out


def best_candidates(cands: List[TtgtCandidate], max_candidates: int) -> List[TtgtCandidate]:
    """Candidates of minimal transposition cost, at most max_candidates
    (selectBestCandidate, tccg/ttgemmt.py:96-114)."""
    if not cands:
        return []
    lo = min(c.cost for c in cands)
    return [c for c in cands if c.cost <= lo * (1 + 1e-12)][:max(1, max_candidates)]


def _prod(xs) -> int:
    p = 1
    for x in xs:
        p *= int(x)
    return p


class DeviceTtgt:
    """One TTGT candidate ready to run on the GPU: transposition plans, the
    GEMM's GETT plan and the workspace for the temporaries."""

    def __init__(self, contraction, cand: TtgtCandidate, prec: int = F32, gemm_plan=None, device=None,
                 tune: bool = False):
        from .contraction import Contraction
        from .runtime import _check_cuda
        from .transpose import DeviceTranspose, plan_transpose
        torch = _check_cuda()
        pb = contraction.problem
        self.contraction, self.cand, self.prec = contraction, cand, prec
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.dtype = torch.float64 if prec == F64 else torch.float32
        self.esz = esz = 8 if prec == F64 else 4
        sizes = {i.label: i.size for T in (pb.opA, pb.opB, pb.C) for i in T.indices}
        X, Y = (pb.opB, pb.opA) if cand.swap else (pb.opA, pb.opB)
        self.x_is_opB = bool(cand.swap)
        # workspace: A~, B~, C~ as needed, each aligned
        offs, total = {}, 0
        for key, need, labels in (("A", cand.transA, cand.idxA), ("B", cand.transB, cand.idxB),
                                  ("C", cand.transC, cand.idxC)):
            if need:
                offs[key] = total
                total += -(-_prod(sizes[l] for l in labels) * esz // _ALIGN) * _ALIGN
        self.offsets = offs
        self.workspace = torch.empty(max(total, _ALIGN), dtype=torch.uint8, device=dev)
        self.tA = DeviceTranspose(plan_transpose(X.labels(), cand.idxA, sizes, elem_bytes=esz), dev) if cand.transA else None
        self.tB = DeviceTranspose(plan_transpose(Y.labels(), cand.idxB, sizes, elem_bytes=esz), dev) if cand.transB else None
        self.tC = (DeviceTranspose(plan_transpose(cand.idxC, pb.C.labels(), sizes, elem_bytes=esz), dev)
                   if cand.transC else None)
        self.gemm = Contraction.from_labels(cand.idxC, cand.idxA, cand.idxB, sizes)
        if tune and gemm_plan is None:
            # the reference's GEMM is a vendor BLAS call; here the GEMM's GETT plans are timed
            self.gemm.tune(prec)
        self.gemm_plan = gemm_plan if gemm_plan is not None else self.gemm.plan(prec)
        self.gemm_dp = self.gemm.device_plan(plan=self.gemm_plan)
        self._sizes = sizes

    def workspace_bytes(self) -> int:
        return int(self.workspace.numel())

    def _view(self, key: str, labels: Sequence[str]):
        n = _prod(self._sizes[l] for l in labels)
        raw = self.workspace[self.offsets[key]: self.offsets[key] + n * self.esz]
        return raw.view(self.dtype).view([self._sizes[l] for l in reversed(list(labels))])

    def run(self, A, B, C, alpha: Optional[float] = None, beta: Optional[float] = None, stream=None) -> None:
        """C <- alpha * contract(A, B) + beta * C with A, B, C in the caller's
        layout (torch shapes reversed(indices) of the .tccg tensors)."""
        c = self.contraction
        alpha = c.alpha if alpha is None else alpha
        beta = c.beta if beta is None else beta
        pb = c.problem
        opA, opB = (B, A) if pb.swapped else (A, B)
        X, Y = (opB, opA) if self.x_is_opB else (opA, opB)
        cand = self.cand
        if self.tA is not None:
            xt = self._view("A", cand.idxA)
            self.tA.run(X.reshape(-1), xt.reshape(-1), 1.0, 0.0, stream)
            X = xt
        if self.tB is not None:
            yt = self._view("B", cand.idxB)
            self.tB.run(Y.reshape(-1), yt.reshape(-1), 1.0, 0.0, stream)
            Y = yt
        gemm_shapes = self.gemm.shapes()
        Xg, Yg = X.reshape(gemm_shapes[0]), Y.reshape(gemm_shapes[1])
        if self.tC is not None:
            ct = self._view("C", cand.idxC)
            self.gemm_dp.run(Xg, Yg, ct.reshape(gemm_shapes[2]), alpha, 0.0, stream=stream)
            self.tC.run(ct.reshape(-1), C.reshape(-1), 1.0, beta, stream)
        else:
            self.gemm_dp.run(Xg, Yg, C.reshape(gemm_shapes[2]), alpha, beta, stream=stream)
