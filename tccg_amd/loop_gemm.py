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

On the GPU every instance of the GEMM for one combination of the contracted
loop indices is one problem of a single batched launch (tccg_gett_batch_f32,
the persistent batched GETT): all problems share one set of offset tables
over (m, n, k), taken with the full tensors' strides, and differ only in
their base pointers.  The contracted loops are consecutive launches.
"""
from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .plan import FREE1, FREE4, INVALID, KDIM1, KDIM4, _vector_ok_free, _vector_ok_k

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


def _prod(xs) -> int:
    p = 1
    for x in xs:
        p *= int(x)
    return p


def _table(extent: int, stride: int, pad_to: int) -> np.ndarray:
    total = -(-extent // pad_to) * pad_to
    out = np.full(total, INVALID, dtype=np.uint32)
    out[:extent] = (np.arange(extent, dtype=np.int64) * stride).astype(np.uint32)
    return out


@dataclass
class LogLayout:
    """Everything the batched launches of one LoG candidate need, host side."""
    shape: int                                # batch tile shape id (tccg_amd.batch.BATCH_SHAPES)
    BM: int
    BN: int
    BK: int
    tables: Tuple[np.ndarray, ...]            # offAm, offAk, offBn, offBk, offCm, offCn, ordA, ordB
    modeA: int
    modeB: int
    free_bases: List[Tuple[int, int, int]]    # element offsets into (opA, opB, C), one per problem
    contracted_bases: List[Tuple[int, int, int]]   # one per launch

    @property
    def tiles(self) -> Tuple[int, int, int]:
        return self.tables[0].size // self.BM, self.tables[2].size // self.BN, self.tables[1].size // self.BK


def log_layout(pb, cand: LogCandidate, shape: Optional[int] = None) -> LogLayout:
    """Shared GEMM tables over (m, n, k) with the full tensors' strides, load
    modes, and the base offsets of every free and contracted loop combination."""
    from .batch import BATCH_SHAPES
    from .plan import F32, TILES
    AA, BB, CC = pb.opA, pb.opB, pb.C
    has = lambda T, l: l in T.labels()
    stride = lambda T, l: T.stride_of_label(l) if (l != DUMMY and has(T, l)) else 0
    ext = lambda l: 1 if l == DUMMY else next(i.size for X in (AA, BB, CC) for i in X.indices if i.label == l)
    Mg, Ng, Kg = ext(cand.m), ext(cand.n), ext(cand.k)
    if shape is None:
        shape = 1 if Mg * Ng >= 256 * 256 else 0
    BM, BN, BK = TILES[F32][BATCH_SHAPES[shape]]
    offAm, offCm = _table(Mg, stride(AA, cand.m), BM), _table(Mg, stride(CC, cand.m), BM)
    offBn, offCn = _table(Ng, stride(BB, cand.n), BN), _table(Ng, stride(CC, cand.n), BN)
    offAk, offBk = _table(Kg, stride(AA, cand.k), BK), _table(Kg, stride(BB, cand.k), BK)
    ordA, ordB = np.arange(BM, dtype=np.int32), np.arange(BN, dtype=np.int32)
    a0, b0 = AA.labels()[0], BB.labels()[0]
    # the stride-one index of A is m or k, that of B n or k (log_candidates)
    if a0 == cand.m:
        modeA = FREE4 if _vector_ok_free(offAm, ordA, BM) else FREE1
    else:
        modeA = KDIM4 if _vector_ok_k(offAk) else KDIM1
    if b0 == cand.n:
        modeB = FREE4 if _vector_ok_free(offBn, ordB, BN) else FREE1
    else:
        modeB = KDIM4 if _vector_ok_k(offBk) else KDIM1

    def bases(labels):
        combos = [(0, 0, 0)]
        for l in labels:
            sa, sb, sc = stride(AA, l), stride(BB, l), stride(CC, l)
            combos = [(x + i * sa, y + i * sb, z + i * sc) for i in range(ext(l)) for (x, y, z) in combos]
        return combos

    free = ([cand.batch] if cand.batch != DUMMY else []) + list(cand.free_order)
    return LogLayout(shape, BM, BN, BK, (offAm, offAk, offBn, offBk, offCm, offCn, ordA, ordB), modeA, modeB,
                     bases(free), bases(list(cand.contracted_order)))


# tccg_batch_problem (include/tccg_gfx950.h) as a numpy record
PROBLEM_DTYPE = np.dtype([(n, "<u8") for n in ("A", "B", "C", "offAm", "offAk", "offBn", "offBk", "offCm", "offCn",
                                               "ordA", "ordB")]
                         + [(n, "<i4") for n in ("tilesM", "tilesN", "kTiles", "variant")]
                         + [("alpha", "<f4"), ("beta", "<f4")])


def problem_records(lay: LogLayout, ptrA: int, ptrB: int, ptrC: int, tables: Sequence[int], alpha: float,
                    beta: float) -> List[np.ndarray]:
    """One record array per launch (combination of the contracted loops):
    problem i of launch j multiplies opA + free_bases[i] + contracted_bases[j]
    by opB + ... into C + free_bases[i], FP32 element offsets; beta on the
    first launch, 1 on the others."""
    tilesM, tilesN, kTiles = lay.tiles
    fb = np.array(lay.free_bases, dtype=np.uint64).reshape(-1, 3)
    rec = np.zeros(fb.shape[0], dtype=PROBLEM_DTYPE)
    for n, v in zip(("offAm", "offAk", "offBn", "offBk", "offCm", "offCn", "ordA", "ordB"), tables):
        rec[n] = v
    rec["tilesM"], rec["tilesN"], rec["kTiles"] = tilesM, tilesN, kTiles
    rec["variant"] = lay.shape * 16 + lay.modeA * 4 + lay.modeB
    rec["alpha"] = alpha
    rec["C"] = np.uint64(ptrC) + np.uint64(4) * fb[:, 2]
    out = []
    for j, (ca, cb, _) in enumerate(lay.contracted_bases):
        rec["A"] = np.uint64(ptrA) + np.uint64(4) * (fb[:, 0] + np.uint64(ca))
        rec["B"] = np.uint64(ptrB) + np.uint64(4) * (fb[:, 1] + np.uint64(cb))
        rec["beta"] = beta if j == 0 else 1.0
        out.append(rec.copy())
    return out


class DeviceLoG:
    """One LoG candidate bound to fixed FP32 tensors: the shared GEMM tables,
    one problem table per combination of the contracted loops (the first with
    beta, the others accumulating), and the work list of the batched launches."""

    def __init__(self, contraction, cand: LogCandidate, A, B, C, alpha: Optional[float] = None,
                 beta: Optional[float] = None, shape: Optional[int] = None):
        import torch
        from .batch import _CProblem
        from .runtime import load_library
        pb = contraction.problem
        self.contraction, self.cand = contraction, cand
        alpha = contraction.alpha if alpha is None else alpha
        beta = contraction.beta if beta is None else beta
        for name, t, T in (("A", A, contraction.A0), ("B", B, contraction.B0), ("C", C, contraction.C0)):
            if (t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != T.size() or t.device.type != "cuda"
                    or t.device != C.device):
                raise ValueError("LoG: %s must be a contiguous float32 tensor of %d elements on C's GPU"
                                 % (name, T.size()))
        opA, opB = (B, A) if pb.swapped else (A, B)
        lay = log_layout(pb, cand, shape)
        self.layout = lay
        chunks, offs, pos = [], [], 0
        for arr in lay.tables:
            a = np.ascontiguousarray(arr).view(np.int32)
            offs.append(pos)
            chunks.append(a)
            pos += a.size
            pad = (-pos) % 4   # every table 16-byte aligned
            if pad:
                chunks.append(np.zeros(pad, np.int32))
                pos += pad
        dev = C.device
        self.tables = torch.from_numpy(np.concatenate(chunks)).to(dev)
        tptr = [self.tables.data_ptr() + 4 * o for o in offs]
        if PROBLEM_DTYPE.itemsize != ctypes.sizeof(_CProblem):
            raise RuntimeError("LoG problem record layout mismatch")
        tilesM, tilesN, _ = lay.tiles
        nprob = len(lay.free_bases)
        work = np.array([[p, tm, tn, 0] for p in range(nprob) for tm in range(tilesM) for tn in range(tilesN)],
                        dtype=np.int32)
        self.work = torch.from_numpy(work).to(dev)
        self.nwork = int(work.shape[0])
        self.grid = min(self.nwork, 512)
        self.launches = [torch.from_numpy(r.view(np.uint8).copy()).to(dev)
                         for r in problem_records(lay, opA.data_ptr(), opB.data_ptr(), C.data_ptr(), tptr,
                                                  alpha, beta)]
        self._tensors = (A, B, C)
        self.lib = load_library()

    def launch(self, stream=None) -> None:
        import torch
        from .runtime import GettRuntimeError
        dev = self.tables.device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        with torch.cuda.device(dev):
            for probs in self.launches:
                rc = self.lib.tccg_gett_batch_f32(ctypes.c_void_p(probs.data_ptr()),
                                                  ctypes.c_void_p(self.work.data_ptr()), self.nwork, self.grid,
                                                  ctypes.c_void_p(s.cuda_stream))
                if rc != 0:
                    raise GettRuntimeError("LoG launch failed: %s" % self.lib.tccg_last_error().decode())

    def run(self, A=None, B=None, C=None, alpha=None, beta=None, stream=None) -> None:
        """Launch on the bound tensors (the arguments exist for the timing
        helpers' calling convention and must be the bound tensors)."""
        for t, b in zip((A, B, C), self._tensors):
            if t is not None and t.data_ptr() != b.data_ptr():
                raise ValueError("LoG runs on the tensors it was bound to")
        self.launch(stream)
