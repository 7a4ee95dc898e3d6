"""Grouped packed GETT: many contractions of one precision and one packed tile
shape in two launches (tccg_gett_group, include/tccg_gfx950.h): one pack
launch over the 64 x 32 blocks of every distinct operand, then one launch of
the packed kernel's one-tile walk with a workgroup per C tile of every
contraction.

Operands are packed once even when several contractions read them (the
loop-over-GEMM method's instances that differ only in an index the operand
lacks): `operand()` returns the index of an operand already added under the
same key.  The work list puts the tiles of contractions that share operands
on one XCD (workgroup i runs on XCD i mod 8) and, inside each XCD's share,
the longest k ranges first.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .plan import BF16, BF16X3, F32, F64, INVALID, KDIM1, KDIM4, PACKED_TAIL_BYTES, PACKED_TILES, _ELEM_BYTES, _PLANES

GROUP_PRECS = (F32, F64, BF16X3, BF16)
PACK_TF, PACK_TK = 64, 32     # block of the pack kernel (free x k)


class _COperand(ctypes.Structure):
    _fields_ = [("X", ctypes.c_void_p), ("offF", ctypes.c_void_p), ("offK", ctypes.c_void_p),
                ("out", ctypes.c_uint64), ("fBlocks", ctypes.c_int32), ("kTiles", ctypes.c_int32),
                ("kfast", ctypes.c_int32), ("isB", ctypes.c_int32)]


class _CProblem(ctypes.Structure):
    _fields_ = [("C", ctypes.c_void_p), ("offCm", ctypes.c_void_p), ("offCn", ctypes.c_void_p),
                ("packedA", ctypes.c_uint64), ("packedB", ctypes.c_uint64),
                ("tilesM", ctypes.c_int32), ("tilesN", ctypes.c_int32), ("kTiles", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("alpha", ctypes.c_double), ("beta", ctypes.c_double)]


GROUP_M = 8


def _walk(tilesM: int, tilesN: int) -> List[Tuple[int, int]]:
    """(tm, tn) of logical tiles 0, 1, ...: groups of GROUP_M tile rows walked
    column by column (tile_coords, gett_core.hpp)."""
    out = []
    for first in range(0, tilesM, GROUP_M):
        gsz = min(tilesM - first, GROUP_M)
        for tn in range(tilesN):
            for r in range(gsz):
                out.append((first + r, tn))
    return out


def kfast(mode: int) -> int:
    """Pack read order of a load mode: along k for the KDIM modes."""
    return 1 if mode in (KDIM4, KDIM1) else 0


class GettGroup:
    """Host builder and launcher of one grouped packed launch.

    add operands with operand(), contractions with problem(), then bind() to
    a device (uploads the tables, records and work list and allocates the
    workspace) and launch() as often as needed; the tensors' pointers are
    baked in."""

    def __init__(self, prec: int, tile: int, order: str = "xcd"):
        if order not in ("xcd", "lpt"):
            raise ValueError("work order is 'xcd' or 'lpt'")
        self.order = order
        if prec not in GROUP_PRECS:
            raise ValueError("the grouped GETT runs FP32, FP64, split-BF16 and BF16")
        if not 0 <= tile < len(PACKED_TILES[prec]):
            raise ValueError("invalid packed tile id %d" % tile)
        self.prec, self.tile = prec, tile
        self.BM, self.BN, self.BK = PACKED_TILES[prec][tile]
        self._tables: List[np.ndarray] = []
        self._table_ids: Dict[int, int] = {}
        self._ops: List[Tuple[int, int, int, int, int, int]] = []   # ptr, tF, tK, kfast, isB, Fp/Kp key
        self._op_keys: Dict[object, int] = {}
        self._probs: List[Tuple] = []
        self._bound = None

    # -- building ---------------------------------------------------------
    def _table(self, arr: np.ndarray) -> int:
        key = id(arr)
        if key not in self._table_ids:
            a = np.ascontiguousarray(arr, dtype=np.uint32)
            self._table_ids[key] = len(self._tables)
            self._tables.append(a)
        return self._table_ids[key]

    def operand(self, ptr: int, offF: np.ndarray, offK: np.ndarray, kfast_: int, is_b: bool,
                key: Optional[object] = None) -> int:
        """Index of the operand (ptr, tables): rows offF (a multiple of 64 and
        of BM, or BN for B operands), k positions offK (a multiple of BK and
        of 32), packed along k when kfast_."""
        F = self.BN if is_b else self.BM
        if offF.size % PACK_TF or offF.size % F:
            raise ValueError("free table of %d entries: not a multiple of 64 and %d" % (offF.size, F))
        if offK.size % self.BK or offK.size % PACK_TK or offK.size == 0:
            raise ValueError("k table of %d entries: not a multiple of %d and 32" % (offK.size, self.BK))
        k = key if key is not None else (int(ptr), id(offF), id(offK), int(kfast_), bool(is_b))
        if k in self._op_keys:
            return self._op_keys[k]
        self._ops.append((int(ptr), self._table(offF), self._table(offK), int(kfast_), int(bool(is_b)),
                          offF.size, offK.size))
        self._op_keys[k] = len(self._ops) - 1
        return self._op_keys[k]

    def problem(self, a: int, b: int, C_ptr: int, offCm: np.ndarray, offCn: np.ndarray, alpha: float,
                beta: float, locality: Optional[object] = None) -> int:
        """C (tables offCm over A's rows, offCn over B's rows) = alpha * A B +
        beta * C for operands a (rows of BM) and b (rows of BN)."""
        oa, ob = self._ops[a], self._ops[b]
        if oa[4] or not ob[4]:
            raise ValueError("operand a must be an A operand and b a B operand")
        if oa[6] != ob[6]:
            raise ValueError("operands with different k ranges")
        if offCm.size != oa[5] or offCn.size != ob[5]:
            raise ValueError("C tables must match the operands' rows")
        self._probs.append((a, b, int(C_ptr), self._table(offCm), self._table(offCn), float(alpha), float(beta),
                            locality if locality is not None else (a, b)))
        return len(self._probs) - 1

    def plane_bytes(self, op) -> int:
        return op[5] * op[6] * _PLANES[self.prec] * _ELEM_BYTES[self.prec]

    def workspace_bytes(self) -> int:
        total = 0
        for op in self._ops:
            total += -(-self.plane_bytes(op) // 256) * 256
        return total + PACKED_TAIL_BYTES

    def work_list(self) -> np.ndarray:
        """(problem, tm, tn) per C tile.  Inside a problem the tiles follow the
        single-problem kernel's walk: groups of GROUP_M tile rows, column by
        column (tile_coords in gett_core.hpp).  Order "lpt": longest k range
        first; a problem's tiles take consecutive positions, and the
        positions that run on one XCD (workgroup i runs on XCD i mod 8) take
        a contiguous range of its walk, as xcd_chunked gives the
        single-problem kernel.  Order "xcd": the problems split by locality
        key into 8 runs of about equal work, item j of run x at position
        8 j + x, longest k range first inside each run; positions past the
        end of a shorter run hold problem -1, which the kernel skips."""
        probs = []
        for j, (a, b, _, _, _, _, _, loc) in enumerate(self._probs):
            oa, ob = self._ops[a], self._ops[b]
            probs.append((loc, oa[6] // self.BK, j, oa[5] // self.BM, ob[5] // self.BN))
        if not probs:
            return np.zeros((0, 4), np.int32)
        if self.order == "lpt":
            out, pos = [], 0
            for loc, kt, j, tM, tN in sorted(probs, key=lambda q: (-q[1], q[2])):
                n = tM * tN
                walk = _walk(tM, tN)
                res = (pos + np.arange(n)) % 8
                logical = np.empty(n, dtype=np.int64)
                start = 0
                for x in range(8):
                    sel = np.nonzero(res == (pos + x) % 8)[0]
                    logical[sel] = start + np.arange(sel.size)
                    start += sel.size
                out.extend((j, int(walk[l][0]), int(walk[l][1])) for l in logical)
                pos += n
            return np.array([[j, tm, tn, 0] for j, tm, tn in out], dtype=np.int32)
        items = []
        for loc, kt, j, tM, tN in probs:
            for tm, tn in _walk(tM, tN):
                items.append((loc, kt, j, tm, tn))
        total = sum(it[1] for it in items)
        runs, acc, cur = [[] for _ in range(8)], 0, 0
        for it in sorted(items, key=lambda it: (repr(it[0]), it[2])):   # stable: walk order inside a problem
            if cur < 7 and acc >= total * (cur + 1) / 8:
                cur += 1
            runs[cur].append(it)
            acc += it[1]
        n = max(len(r) for r in runs)
        out = np.full((8 * n, 4), 0, dtype=np.int32)
        out[:, 0] = -1
        for x, r in enumerate(runs):
            r.sort(key=lambda it: -it[1])
            for j, it in enumerate(r):
                out[8 * j + x, :3] = it[2:]
        return out

    # -- records ----------------------------------------------------------
    def records(self, table_base: int):
        """Host images of everything the launch reads, with the tables placed
        at device address table_base: (tables as one int32 array, operand
        records, pack jobs (n x 4 int32), problem records, work list,
        workspace bytes)."""
        chunks, offs, pos = [], [], 0
        for a in self._tables:
            v = a.view(np.int32)
            offs.append(pos)
            chunks.append(v)
            pos += v.size
            pad = (-pos) % 4    # every table 16-byte aligned
            if pad:
                chunks.append(np.zeros(pad, np.int32))
                pos += pad
        tables = np.concatenate(chunks) if chunks else np.zeros(4, np.int32)
        tptr = [table_base + 4 * o for o in offs]
        ops = (_COperand * max(1, len(self._ops)))()
        out_off, jobs, pos = [], [], 0
        for i, (ptr, tF, tK, kf, isb, Fp, Kp) in enumerate(self._ops):
            o = ops[i]
            o.X, o.offF, o.offK, o.out = ptr, tptr[tF], tptr[tK], pos
            o.fBlocks, o.kTiles, o.kfast, o.isB = Fp // PACK_TF, Kp // self.BK, kf, isb
            out_off.append(pos)
            for k in range(Kp // PACK_TK):
                for f in range(Fp // PACK_TF):
                    jobs.append((i, f, k, 0))
            pos += -(-self.plane_bytes(self._ops[i]) // 256) * 256
        probs = (_CProblem * max(1, len(self._probs)))()
        for j, (a, b, cptr, tm_, tn_, alpha, beta, _) in enumerate(self._probs):
            q, oa, ob = probs[j], self._ops[a], self._ops[b]
            q.C, q.offCm, q.offCn = cptr, tptr[tm_], tptr[tn_]
            q.packedA, q.packedB = out_off[a], out_off[b]
            q.tilesM, q.tilesN, q.kTiles = oa[5] // self.BM, ob[5] // self.BN, oa[6] // self.BK
            q.alpha, q.beta = alpha, beta
        jobs = np.array(jobs, dtype=np.int32).reshape(-1, 4)
        return tables, ops, jobs, probs, self.work_list(), pos + PACKED_TAIL_BYTES

    # -- device -----------------------------------------------------------
    def bind(self, device, workspace=None):
        import torch
        from .runtime import load_library
        dev = torch.device(device)
        n32 = sum(a.size + (-a.size) % 4 for a in self._tables) or 4
        tables = torch.empty(n32, dtype=torch.int32, device=dev)
        host_tables, ops, jobs, probs, work, need = self.records(tables.data_ptr())
        tables.copy_(torch.from_numpy(host_tables))
        if workspace is None:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
        elif workspace.dtype != torch.uint8 or workspace.numel() < need or workspace.device != dev:
            raise ValueError("group workspace must be a uint8 tensor of >= %d bytes on %s" % (need, dev))
        up = lambda arr: torch.from_numpy(np.frombuffer(bytes(arr), dtype=np.uint8).copy()).to(dev)
        self._bound = dict(tables=tables, ops=up(ops), nops=len(self._ops), jobs=torch.from_numpy(jobs).to(dev),
                           njobs=int(jobs.shape[0]), probs=up(probs), nprobs=len(self._probs),
                           work=torch.from_numpy(work).to(dev), nwork=int(work.shape[0]), ws=workspace,
                           lib=load_library())
        return self

    def launch(self, stream=None) -> None:
        import torch
        from .runtime import GettRuntimeError
        b = self._bound
        if b is None:
            raise RuntimeError("bind() the group to a device first")
        dev = b["ws"].device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        vp = ctypes.c_void_p
        with torch.cuda.device(dev):
            rc = b["lib"].tccg_gett_group(self.prec, self.tile, vp(b["ops"].data_ptr()), b["nops"],
                                          vp(b["jobs"].data_ptr()), b["njobs"], vp(b["probs"].data_ptr()),
                                          b["nprobs"], vp(b["work"].data_ptr()), b["nwork"],
                                          vp(b["ws"].data_ptr()), b["ws"].numel(), vp(s.cuda_stream))
        if rc != 0:
            raise GettRuntimeError("grouped GETT failed: %s" % b["lib"].tccg_last_error().decode())


def group_for_plans(plans: Sequence, As, Bs, Cs, alphas: Sequence[float], betas: Sequence[float],
                    order: str = "xcd") -> GettGroup:
    """One group over packed plans of one precision and tile (built by
    plan.build_packed_plan): each contraction's internal operands (A holds
    C's stride-one index) with the plan's tables."""
    if not plans:
        raise ValueError("empty group")
    prec, tile = plans[0].prec, plans[0].tile
    g = GettGroup(prec, tile, order)
    for p, A, B, C, al, be in zip(plans, As, Bs, Cs, alphas, betas):
        if p.prec != prec or p.tile != tile or not p.packed or p.ksplit > 1 or p.walk != 0:
            raise ValueError("group plans must be one-tile packed plans of one precision and tile")
        opA, opB = (B, A) if p.problem.swapped else (A, B)
        a = g.operand(opA.data_ptr(), p.offAm, p.offAk, kfast(p.modeA), False)
        b = g.operand(opB.data_ptr(), p.offBn, p.offBk, kfast(p.modeB), True)
        g.problem(a, b, C.data_ptr(), p.offCm, p.offCn, al, be)
    return g
