"""Batched GETT: hundreds of uneven FP32 contractions in one persistent launch.

Each contraction gets a GETT plan (tables as for a single launch); the plans'
tables are concatenated into one device buffer, a problem table
(tccg_batch_problem, include/tccg_gfx950.h) points into it, and a work list
enumerates every C tile of every problem, longest k-loop first (LPT order), so
that the persistent grid's round-robin walk balances the load.  The whole
batch is one kernel launch, which can be captured in a HIP graph.

Tile shapes are chosen per index pattern: `autotune_patterns()` times each
pattern's problems with each batch tile shape (the GPU analogue of the
reference's timed search, tccg/tccg.py:278-430) and keeps the faster one.
"""
from __future__ import annotations

import ctypes
from collections import defaultdict
from typing import Dict, List, Optional, Sequence

import numpy as np

from .contraction import Contraction
from .plan import F32, PACKED_TILES, TILES, build_packed_plan, build_plan

# batch shape id -> tile id of TILES[F32] with the same (BM, BN, BK)
BATCH_SHAPES = {0: TILES[F32].index((64, 64, 32)), 1: TILES[F32].index((128, 128, 32))}


class _CProblem(ctypes.Structure):
    _fields_ = [("A", ctypes.c_void_p), ("B", ctypes.c_void_p), ("C", ctypes.c_void_p),
                ("offAm", ctypes.c_void_p), ("offAk", ctypes.c_void_p), ("offBn", ctypes.c_void_p),
                ("offBk", ctypes.c_void_p), ("offCm", ctypes.c_void_p), ("offCn", ctypes.c_void_p),
                ("ordA", ctypes.c_void_p), ("ordB", ctypes.c_void_p),
                ("tilesM", ctypes.c_int32), ("tilesN", ctypes.c_int32), ("kTiles", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("alpha", ctypes.c_float), ("beta", ctypes.c_float)]


def pattern_key(c: Contraction) -> str:
    return "%s=%s*%s" % (c.C0.index_str(), c.A0.index_str(), c.B0.index_str())


# packed tile of the problems that run on the packed path (FP32 128 x 128 x 16)
LARGE_TILE = PACKED_TILES[F32].index((128, 128, 16))


class GettBatch:
    """Fixed set of contractions bound to fixed GPU tensors, launched together.

    packed_min_tiles: problems whose packed plan (128 x 128 tiles) has at least
    this many C tiles, enough to fill the GPU on their own, run one after the
    other on the packed path (pack kernels + DMA-ring GETT, sharing one
    workspace) on the same stream, ahead of the single persistent launch that
    runs all other problems.  None: every problem in the persistent launch."""

    def __init__(self, contractions: Sequence[Contraction], shapes: Optional[Dict[str, int]] = None,
                 alphas: Optional[Sequence[float]] = None, betas: Optional[Sequence[float]] = None,
                 packed_min_tiles: Optional[int] = None, group_tile: Optional[int] = None,
                 group_order: str = "xcd"):
        self.contractions = list(contractions)
        self.shapes = dict(shapes or {})
        self.alphas = list(alphas) if alphas is not None else [c.alpha for c in self.contractions]
        self.betas = list(betas) if betas is not None else [c.beta for c in self.contractions]
        self.plans = []
        self.large: Dict[int, object] = {}      # problem index -> packed plan
        # group_tile: every problem on the grouped packed path (one pack
        # launch and one GETT launch, tccg_amd.group) with this FP32 packed
        # tile and work order
        self.group_tile = group_tile
        self.group_order = group_order
        if group_tile is not None:
            self.plans = [build_packed_plan(c.problem, F32, group_tile) for c in self.contractions]
            self._bound = None
            return
        for i, c in enumerate(self.contractions):
            if packed_min_tiles is not None:
                pp = build_packed_plan(c.problem, F32, LARGE_TILE)
                if pp.ntiles() >= packed_min_tiles:
                    self.large[i] = pp
            sh = self.shapes.get(pattern_key(c), self._default_shape(c))
            self.plans.append(build_plan(c.problem, F32, BATCH_SHAPES[sh]))
        self._bound = None

    @staticmethod
    def _default_shape(c: Contraction) -> int:
        return 1 if c.problem.M * c.problem.N >= 256 * 256 else 0

    def flops(self) -> float:
        return sum(c.flops(b) for c, b in zip(self.contractions, self.betas))

    def bind(self, As, Bs, Cs, device=None, grid: Optional[int] = None):
        """Build the device problem/work tables for these tensors (their
        pointers are baked in; reuse the same tensors for every launch)."""
        import torch
        from .runtime import load_library
        from .runtime import DevicePlan
        lib = load_library()
        from .runtime import cuda_device
        dev = cuda_device(torch, device if device is not None else As[0].device)
        self._check_tensors(As, Bs, Cs, dev)
        if self.group_tile is not None:
            from .group import group_for_plans
            grp = group_for_plans(self.plans, As, Bs, Cs, self.alphas, self.betas, self.group_order).bind(dev)
            self._bound = dict(group=grp, tensors=(list(As), list(Bs), list(Cs)))
            return self
        # problems on the packed path: one shared workspace (they run one after the other)
        large = []
        if self.large:
            ws = max(p.workspace_bytes() for p in self.large.values())
            shared = torch.empty(ws, dtype=torch.uint8, device=dev)
            for i, p in self.large.items():
                large.append((DevicePlan(p, device=dev, workspace=shared), i))
        small = [i for i in range(len(self.plans)) if i not in self.large]
        # all tables of all problems in one int32 buffer
        chunks, offsets, pos = [], [], 0
        for p in (self.plans[i] for i in small):
            offs = []
            for arr in (p.offAm, p.offAk, p.offBn, p.offBk, p.offCm, p.offCn, p.ordA, p.ordB):
                a = np.ascontiguousarray(arr).view(np.int32)
                offs.append(pos)
                chunks.append(a)
                pos += a.size
                pad = (-pos) % 4   # keep every table 16-byte aligned
                if pad:
                    chunks.append(np.zeros(pad, np.int32))
                    pos += pad
            offsets.append(offs)
        tables = torch.from_numpy(np.concatenate(chunks) if chunks else np.zeros(4, np.int32)).to(dev)
        base = tables.data_ptr()
        probs = (_CProblem * max(1, len(small)))()
        work = []
        for j in range(len(small)):
            i = small[j]
            p, c, A, B, C = self.plans[i], self.contractions[i], As[i], Bs[i], Cs[i]
            opA, opB = (B, A) if p.problem.swapped else (A, B)
            q = probs[j]
            q.A, q.B, q.C = opA.data_ptr(), opB.data_ptr(), C.data_ptr()
            (q.offAm, q.offAk, q.offBn, q.offBk, q.offCm, q.offCn, q.ordA, q.ordB) = \
                [base + 4 * o for o in offsets[j]]
            q.tilesM, q.tilesN, q.kTiles = p.Mp // p.BM, p.Np // p.BN, p.Kp // p.BK
            sh = [k for k, v in BATCH_SHAPES.items() if v == p.tile][0]
            q.variant = sh * 16 + p.modeA * 4 + p.modeB
            q.alpha, q.beta = self.alphas[i], self.betas[i]
            cost = p.BM * p.BN * q.kTiles
            for tm in range(q.tilesM):
                for tn in range(q.tilesN):
                    work.append((cost, j, tm, tn))   # j: row of the problem table
        # longest tiles first: the round-robin walk then approximates LPT
        work.sort(key=lambda w: -w[0])
        wl = np.array([[w[1], w[2], w[3], 0] for w in work], dtype=np.int32).reshape(-1, 4)
        pbytes = np.frombuffer(bytes(probs), dtype=np.uint8)
        dprobs = torch.from_numpy(pbytes.copy()).to(dev)
        dwork = torch.from_numpy(wl).to(dev)
        if grid is None:
            grid = max(1, min(len(work), 256 * 2))
        self._bound = dict(tables=tables, probs=dprobs, work=dwork, nwork=len(work), grid=int(grid),
                           tensors=(list(As), list(Bs), list(Cs)), lib=lib, large=large)
        return self

    def _check_tensors(self, As, Bs, Cs, dev) -> None:
        import torch
        if not len(As) == len(Bs) == len(Cs) == len(self.contractions):
            raise ValueError("one A, B and C tensor per contraction")
        for i, c in enumerate(self.contractions):
            for name, t, T in (("A", As[i], c.A0), ("B", Bs[i], c.B0), ("C", Cs[i], c.C0)):
                if t.dtype != torch.float32 or not t.is_contiguous() or t.numel() != T.size() or t.device != dev:
                    raise ValueError("problem %d: %s must be a contiguous float32 tensor of %d elements on %s"
                                     % (i, name, T.size(), dev))

    def launches(self) -> int:
        """Kernel launches of one batch step."""
        b = self._bound
        if b is None:
            raise RuntimeError("bind() the batch to tensors first")
        if "group" in b:
            return 2
        return 3 * len(b["large"]) + (1 if b["nwork"] else 0)

    def launch(self, stream=None) -> None:
        import torch
        from .runtime import GettRuntimeError
        b = self._bound
        if b is None:
            raise RuntimeError("bind() the batch to tensors first")
        if "group" in b:
            b["group"].launch(stream)
            return
        dev = b["work"].device
        s = stream if stream is not None else torch.cuda.current_stream(dev)
        As, Bs, Cs = b["tensors"]
        for dp, i in b["large"]:   # one after the other: they share the workspace
            dp.run(As[i], Bs[i], Cs[i], self.alphas[i], self.betas[i], stream=s)
        if b["nwork"] == 0:
            return
        with torch.cuda.device(dev):
            rc = b["lib"].tccg_gett_batch_f32(ctypes.c_void_p(b["probs"].data_ptr()),
                                              ctypes.c_void_p(b["work"].data_ptr()), b["nwork"], b["grid"],
                                              ctypes.c_void_p(s.cuda_stream))
        if rc != 0:
            raise GettRuntimeError("batched GETT failed: %s" % b["lib"].tccg_last_error().decode())

    def capture(self):
        """A torch.cuda.CUDAGraph (HIP graph) holding the one batched launch."""
        import torch
        self.launch()              # first launch outside capture sets kernel attributes
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.launch()
        torch.cuda.current_stream().wait_stream(s)
        return g


def random_batch(n: int, seed: int, patterns=None, lo: int = 8, hi: int = 96, step: int = 8):
    """n contractions drawn uniformly from `patterns`, every extent uniform over
    multiples of `step` in [lo, hi] (numpy default_rng(seed))."""
    from .workloads import BATCH_PATTERNS
    patterns = patterns or BATCH_PATTERNS
    rng = np.random.default_rng(seed)
    choices = np.arange(lo, hi + 1, step)
    out = []
    for _ in range(n):
        c_l, a_l, b_l = patterns[int(rng.integers(len(patterns)))]
        labels = sorted(set(c_l + a_l + b_l))
        sizes = {l: int(rng.choice(choices)) for l in labels}
        out.append(Contraction.from_labels(c_l, a_l, b_l, sizes))
    return out


def autotune_patterns(contractions: Sequence[Contraction], As, Bs, Cs, reps: int = 3, verbose: bool = False):
    """Per index pattern, the batch tile shape with the shortest batched launch
    over that pattern's problems."""
    import torch
    groups = defaultdict(list)
    for i, c in enumerate(contractions):
        groups[pattern_key(c)].append(i)
    best = {}
    for key, idx in groups.items():
        times = {}
        for sh in BATCH_SHAPES:
            b = GettBatch([contractions[i] for i in idx], {key: sh}).bind(
                [As[i] for i in idx], [Bs[i] for i in idx], [Cs[i] for i in idx])
            b.launch()
            torch.cuda.synchronize()
            t = float("inf")
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                b.launch()
                e1.record()
                e1.synchronize()
                t = min(t, e0.elapsed_time(e1) * 1e-3)
            times[sh] = t
        best[key] = min(times, key=times.get)
        if verbose:
            import sys
            print("  pattern %-14s %3d problems: %s -> shape %d" % (
                key, len(idx), {k: "%.1f us" % (v * 1e6) for k, v in times.items()}, best[key]), file=sys.stderr)
    return best


def group_candidates(contractions: Sequence[Contraction], betas: Optional[Sequence[float]] = None):
    """(tile, order) choices of the grouped launch ranked by the planner's
    model of a grouped launch (plan.group_estimate: every tile of every
    problem sharing the GPU; the gfx950 analogue of the reference's
    model-ranked candidates, tccg/gett.py:987-1150), each tile with both
    work orders."""
    from .plan import group_estimate
    betas = list(betas) if betas is not None else [c.beta for c in contractions]
    est = []
    for t in range(len(PACKED_TILES[F32])):
        est.append((group_estimate([build_packed_plan(c.problem, F32, t) for c in contractions], betas), t))
    est.sort()
    return [(t, order, e) for e, t in est for order in ("xcd", "lpt")]


def autotune_group(contractions: Sequence[Contraction], As, Bs, Cs, top: int = 6, reps: int = 2,
                   verbose: bool = False):
    """Time the `top` model-ranked (tile, work order) choices of the grouped
    launch on the bound tensors (the reference's measure-and-select,
    tccg/tccg.py:278-430) and return the fastest as (tile, order, seconds)."""
    import torch
    best = None
    for tile, order, est in group_candidates(contractions)[:top]:
        b = GettBatch(contractions, group_tile=tile, group_order=order).bind(As, Bs, Cs)
        b.launch()
        torch.cuda.synchronize()
        t = float("inf")
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            b.launch()
            e1.record()
            e1.synchronize()
            t = min(t, e0.elapsed_time(e1) * 1e-3)
        if verbose:
            import sys
            print("  group tile %s order %s: model %.2f ms, measured %.2f ms" % (
                PACKED_TILES[F32][tile], order, est * 1e3, t * 1e3), file=sys.stderr)
        if best is None or t < best[2]:
            best = (tile, order, t)
        del b
    return best
