"""Tensor transposition on the GPU: out = alpha * permute(in) + beta * out.

The gfx950 replacement of the transposition step of TCCG's TTGT method
(reference: the Transpose code block, tccg/transpose.py:74-144, which emits
an HPTT plan; used by tccg/ttgemmt.py:133-336).  The planner fuses indices
that stay adjacent in both tensors, splits the rest into

  X: a prefix of the input's indices (consecutive x = consecutive input elements),
  Y: a prefix of the output's indices (consecutive y = consecutive output elements),
  Z: the remaining indices,

and tabulates separable element offsets over each set
(csrc/transpose.hpp).  When input and output share their stride-one index
the transposition is a copy of contiguous rows.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

INVALID = np.uint32(0xFFFFFFFF)

# (TX, TY) per tile id; id 0 is the row copy (csrc/transpose.hip kTransposeTiles)
TRANSPOSE_TILES: List[Tuple[int, int]] = [(0, 0), (64, 64), (32, 32), (32, 64), (64, 32), (128, 32), (32, 128),
                                          (16, 64), (64, 16)]

_TARGET = 64   # elements a tile side aims for when it gathers several indices


def _strides(labels: Sequence[str], sizes: Dict[str, int]) -> Dict[str, int]:
    out, s = {}, 1
    for l in labels:
        out[l] = s
        s *= sizes[l]
    return out


def fuse_groups(in_labels: Sequence[str], out_labels: Sequence[str]) -> List[List[str]]:
    """Maximal runs of indices that are adjacent, in the same order, in both
    tensors (in input order).  Each run behaves as one index."""
    pos_out = {l: i for i, l in enumerate(out_labels)}
    groups: List[List[str]] = []
    for l in in_labels:
        if groups and pos_out[l] == pos_out[groups[-1][-1]] + 1:
            groups[-1].append(l)
        else:
            groups.append([l])
    return groups


@dataclass
class TransposePlan:
    in_labels: List[str]
    out_labels: List[str]
    sizes: Dict[str, int]
    X: List[List[str]]      # index groups of the x set, input order
    Y: List[List[str]]      # index groups of the y set, output order
    Z: List[List[str]]      # remaining groups, input order
    tile: int
    vec: int
    inX: np.ndarray
    outX: np.ndarray
    inY: np.ndarray
    outY: np.ndarray
    inZ: np.ndarray
    outZ: np.ndarray

    @property
    def Xp(self) -> int:
        return int(self.inX.size) if self.tile else _prod(self.sizes[l] for g in self.X for l in g)

    @property
    def Yp(self) -> int:
        return int(self.inY.size) if self.tile else 1

    @property
    def nZ(self) -> int:
        return int(self.inZ.size)

    def elements(self) -> int:
        return _prod(self.sizes[l] for l in self.in_labels)

    def moved_bytes(self, esz: int, beta: float = 0.0) -> int:
        """Algorithmic HBM bytes: read the input, write the output (and read it when beta != 0)."""
        return esz * self.elements() * (3 if beta != 0 else 2)

    def describe(self) -> str:
        j = lambda gs: "(" + ",".join("".join(g) for g in gs) + ")"
        shape = ("rows" if self.tile == 0 else "%dx%d" % TRANSPOSE_TILES[self.tile]) + " vec=%d" % self.vec
        return "transpose %s -> %s tile=%d %s X=%s Y=%s Z=%s" % (
            ",".join(self.in_labels), ",".join(self.out_labels), self.tile, shape, j(self.X), j(self.Y), j(self.Z))


def _prod(xs) -> int:
    p = 1
    for x in xs:
        p *= int(x)
    return p


def _offsets(groups: List[List[str]], sizes: Dict[str, int], stride: Dict[str, int], pad_to: int) -> np.ndarray:
    """uint32 offsets over the groups' positions, first group fastest (each
    group's labels in their own order, first fastest), padded with INVALID to
    a multiple of pad_to."""
    off = np.zeros(1, dtype=np.int64)
    for g in groups:
        for l in g:
            # the new label varies slowest: position = previous + n_previous * i_l
            off = (off[None, :] + stride[l] * np.arange(sizes[l], dtype=np.int64)[:, None]).reshape(-1)
    n = off.size
    total = -(-n // pad_to) * pad_to if pad_to > 1 else n
    out = np.full(total, INVALID, dtype=np.uint32)
    out[:n] = off.astype(np.uint32)
    return out


def _take(groups: List[List[str]], sizes: Dict[str, int], stop_at, limit: int) -> List[List[str]]:
    """Leading groups (at least one) until one in stop_at or a product >= limit."""
    out, p = [], 1
    for g in groups:
        if (out and p >= limit) or any(l in stop_at for l in g):
            break
        out.append(g)
        p *= _prod(sizes[l] for l in g)
    return out


# Tiles in the order of their measured rates on MI355X (128 x 32 moved 4.2-5.2
# TB/s on every shape timed, the others up to 25 % less); padded positions
# issue no memory access, so a tile is kept while it walks at most 2.5x the
# positions it moves.
_TILE_PREFERENCE = [5, 1, 2, 3, 4, 6, 8, 7]


def _best_tile(nx: int, ny: int) -> int:
    """The first preferred tile that pads the x and y extents at most 2.5x,
    else the tile with the least padding."""
    def ratio(t):
        tx, ty = TRANSPOSE_TILES[t]
        return (-(-nx // tx) * tx) * (-(-ny // ty) * ty) / (nx * ny)
    for t in _TILE_PREFERENCE:
        if ratio(t) <= 2.5:
            return t
    return min(_TILE_PREFERENCE, key=ratio)


def plan_transpose(in_labels: Sequence[str], out_labels: Sequence[str], sizes: Dict[str, int],
                   tile: Optional[int] = None, elem_bytes: int = 4) -> TransposePlan:
    in_labels, out_labels = list(in_labels), list(out_labels)
    if sorted(in_labels) != sorted(out_labels) or len(set(in_labels)) != len(in_labels):
        raise ValueError("a transposition needs the same distinct indices on both sides: %s -> %s"
                         % (in_labels, out_labels))
    if _prod(sizes[l] for l in in_labels) >= 0xFFFFFFFF:
        raise ValueError("tensor exceeds 2^32 - 1 elements (32-bit offsets)")
    sin, sout = _strides(in_labels, sizes), _strides(out_labels, sizes)
    groups = fuse_groups(in_labels, out_labels)
    first_out = out_labels[0]
    g_out = sorted(groups, key=lambda g: out_labels.index(g[0]))
    if groups[0][0] == first_out or tile == 0:
        if groups[0][0] != first_out:
            raise ValueError("the row copy needs input and output to share their stride-one index")
        # Rows of at least 1 KiB are walked in output order (consecutive rows
        # written back to back), shorter ones in input order (read back to
        # back): measured on MI355X, 2.3 KB rows move 7 % faster in output
        # order, 256-288 B rows 2-14 % faster in input order.
        row_bytes = elem_bytes * _prod(sizes[l] for l in groups[0])
        rest = groups[1:] if row_bytes < 1024 else sorted(groups[1:], key=lambda g: out_labels.index(g[0]))
        X, Y, Z = [groups[0]], [], rest
        inZ = _offsets(Z, sizes, sin, 1)
        outZ = _offsets(Z, sizes, sout, 1)
        nx = _prod(sizes[l] for l in X[0])
        v = 16 // elem_bytes
        vec = v if (nx % v == 0 and np.all(inZ.astype(np.int64) % v == 0) and np.all(outZ.astype(np.int64) % v == 0)) else 1
        one = np.zeros(1, dtype=np.uint32)
        return TransposePlan(in_labels, out_labels, dict(sizes), X, Y, Z, 0, vec, one, one, one, one, inZ, outZ)
    X = _take(groups, sizes, {first_out}, _TARGET)
    xl = {l for g in X for l in g}
    Y = _take(g_out, sizes, xl, _TARGET)
    yl = {l for g in Y for l in g}
    Z = [g for g in groups if g[0] not in xl and g[0] not in yl]
    nx = _prod(sizes[l] for g in X for l in g)
    ny = _prod(sizes[l] for g in Y for l in g)
    if tile is None:
        tile = _best_tile(nx, ny)
    if not 1 <= tile < len(TRANSPOSE_TILES):
        raise ValueError("invalid transpose tile %d" % tile)
    TX, TY = TRANSPOSE_TILES[tile]
    # 16-byte accesses when both extents are whole vectors: X is the input's
    # prefix and Y the output's, so every other stride is a multiple of them
    v = 16 // elem_bytes
    vec = v if (nx % v == 0 and ny % v == 0) else 1
    return TransposePlan(in_labels, out_labels, dict(sizes), X, Y, Z, tile, vec,
                         _offsets(X, sizes, sin, TX), _offsets(X, sizes, sout, TX),
                         _offsets(Y, sizes, sin, TY), _offsets(Y, sizes, sout, TY),
                         _offsets(Z, sizes, sin, 1), _offsets(Z, sizes, sout, 1))


class _CTransposePlan(ctypes.Structure):
    _fields_ = [("inX", ctypes.c_void_p), ("outX", ctypes.c_void_p), ("inY", ctypes.c_void_p),
                ("outY", ctypes.c_void_p), ("inZ", ctypes.c_void_p), ("outZ", ctypes.c_void_p),
                ("Xp", ctypes.c_int32), ("Yp", ctypes.c_int32), ("Z", ctypes.c_int32),
                ("tile", ctypes.c_int32), ("vec", ctypes.c_int32)]


class DeviceTranspose:
    """A TransposePlan whose tables live on the GPU; run() launches
    tccg_transpose_f32 / tccg_transpose_f64."""

    def __init__(self, plan: TransposePlan, device=None):
        from .runtime import _check_cuda, cuda_device
        torch = _check_cuda()
        self.plan = plan
        dev = cuda_device(torch, device)
        self.device = dev
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).to(dev)
        self.tables = [up(a) for a in (plan.inX, plan.outX, plan.inY, plan.outY, plan.inZ, plan.outZ)]
        c = _CTransposePlan()
        c.inX, c.outX, c.inY, c.outY, c.inZ, c.outZ = [t.data_ptr() for t in self.tables]
        c.Xp, c.Yp, c.Z = plan.Xp, plan.Yp, plan.nZ
        c.tile, c.vec = plan.tile, plan.vec
        self.cplan = c

    def run(self, inp, out, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        from .runtime import GettRuntimeError, check_tensor, last_error, load_library, _torch
        torch = _torch()
        if not isinstance(inp, torch.Tensor) or inp.dtype not in (torch.float32, torch.float64):
            raise TypeError("transpose: float32 or float64 tensors of one dtype")
        n = self.plan.elements()
        check_tensor("input", inp, inp.dtype, n, self.device)
        check_tensor("output", out, inp.dtype, n, self.device)
        lib = load_library()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        fn = lib.tccg_transpose_f64 if inp.dtype == torch.float64 else lib.tccg_transpose_f32
        with torch.cuda.device(self.device):
            rc = fn(inp.data_ptr(), out.data_ptr(), alpha, beta, ctypes.byref(self.cplan),
                    ctypes.c_void_p(s.cuda_stream))
        if rc != 0:
            raise GettRuntimeError("transpose launch failed: %s" % last_error())


def transpose(x, in_labels: Sequence[str], out_labels: Sequence[str], alpha: float = 1.0, beta: float = 0.0,
              out=None):
    """out[out_labels] = alpha * x[in_labels] + beta * out for a torch tensor x
    in TCCG layout (torch shape reversed(in_labels)); returns out."""
    from .runtime import _torch
    torch = _torch()
    in_labels, out_labels = list(in_labels), list(out_labels)
    sizes = dict(zip(in_labels, reversed(list(x.shape))))
    if out is None:
        out = torch.empty([sizes[l] for l in reversed(out_labels)], dtype=x.dtype, device=x.device)
    plan = plan_transpose(in_labels, out_labels, sizes, elem_bytes=x.element_size())
    DeviceTranspose(plan, x.device).run(x, out, alpha, beta)
    return out
