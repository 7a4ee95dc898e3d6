"""ctypes layer between Python/PyTorch-ROCm and the gfx950 GETT library.

The library (tccg_amd/_lib/libtccg_gfx950.so, built by tccg_amd.build) exports
the C ABI of include/tccg_gfx950.h.  This module uploads a plan's offset
tables to the device once, and launches GETT on torch tensors (device
pointers from Tensor.data_ptr()) on torch's current HIP stream.  There is no
CPU fallback: if the library or a GPU is missing, calls raise.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path
from typing import Optional

import numpy as np

from . import build as _build
from .plan import BF16, BF16X3, F32, F64, FP8, PACKED_TILES, VECTOR_TILES, GettPlan, TILES


class GettRuntimeError(RuntimeError):
    pass


class _CPlan(ctypes.Structure):
    _fields_ = [
        ("offAm", ctypes.c_void_p), ("offAk", ctypes.c_void_p),
        ("offBn", ctypes.c_void_p), ("offBk", ctypes.c_void_p),
        ("offCm", ctypes.c_void_p), ("offCn", ctypes.c_void_p),
        ("ordA", ctypes.c_void_p), ("ordB", ctypes.c_void_p),
        ("Mp", ctypes.c_int32), ("Np", ctypes.c_int32), ("Kp", ctypes.c_int32),
        ("tile", ctypes.c_int32), ("modeA", ctypes.c_int32), ("modeB", ctypes.c_int32),
        ("groupM", ctypes.c_int32), ("grid", ctypes.c_int32),
        ("amax", ctypes.c_void_p), ("walk", ctypes.c_int32),
        ("crun_m", ctypes.c_int32), ("crun_n", ctypes.c_int32), ("cvec", ctypes.c_int32),
        ("ksplit", ctypes.c_int32),
    ]


EXPORTED_SYMBOLS = ["tccg_tile_shape", "tccg_num_tiles", "tccg_gett_f32", "tccg_gett_f64",
                    "tccg_gett_bf16x3", "tccg_gett_fp8", "tccg_gett_bf16", "tccg_amax_f32", "tccg_last_error",
                    "tccg_batch_num_variants", "tccg_batch_variant", "tccg_gett_batch_f32",
                    "tccg_packed_workspace_bytes", "tccg_packed_tile_shape", "tccg_num_packed_tiles",
                    "tccg_gett_packed_f32", "tccg_gett_packed_f64", "tccg_gett_packed_bf16x3",
                    "tccg_gett_packed_bf16", "tccg_gett_packed_fp8", "tccg_gett_prepacked",
                    "tccg_num_vector_tiles", "tccg_vector_tile_shape",
                    "tccg_num_transpose_tiles", "tccg_transpose_tile_shape", "tccg_transpose_f32",
                    "tccg_transpose_f64", "tccg_packed_sw_available", "tccg_gett_packed_noalloc",
                    "tccg_gett_group"]

_PACKED_FN = {F32: "tccg_gett_packed_f32", F64: "tccg_gett_packed_f64", BF16X3: "tccg_gett_packed_bf16x3",
              BF16: "tccg_gett_packed_bf16", FP8: "tccg_gett_packed_fp8"}

_lib = None
_lib_lock = threading.Lock()


def load_library(path: Optional[Path] = None) -> ctypes.CDLL:
    """Load (once) the in-tree GETT library (or the one named by $TCCG_GFX950_LIB);
    builds the in-tree library only if it is absent."""
    global _lib
    with _lib_lock:
        if _lib is not None:
            return _lib
        import os
        env = os.environ.get("TCCG_GFX950_LIB")
        p = Path(path) if path else (Path(env) if env else _build.library_path())
        if not p.exists():
            if not _build.hipcc_available():
                raise GettRuntimeError("GETT library %s is missing and hipcc is unavailable" % p)
            _build.build_library()
        lib = ctypes.CDLL(str(p))
        vp, i32, f32, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_double
        lib.tccg_tile_shape.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        lib.tccg_num_tiles.argtypes = [ctypes.c_int]
        for name, T in (("tccg_gett_f32", f32), ("tccg_gett_bf16x3", f32), ("tccg_gett_fp8", f32),
                        ("tccg_gett_bf16", f32), ("tccg_gett_f64", f64)):
            fn = getattr(lib, name)
            fn.argtypes = [vp, vp, vp, T, T, ctypes.POINTER(_CPlan), vp]
            fn.restype = ctypes.c_int
        for prec, name in _PACKED_FN.items():
            fn = getattr(lib, name)
            T = f64 if prec == F64 else f32
            fn.argtypes = [vp, vp, vp, T, T, ctypes.POINTER(_CPlan), vp, ctypes.c_size_t, vp]
            fn.restype = ctypes.c_int
        lib.tccg_gett_prepacked.argtypes = [ctypes.c_int, vp, f64, f64, ctypes.POINTER(_CPlan), vp, ctypes.c_size_t, vp]
        lib.tccg_gett_prepacked.restype = ctypes.c_int
        lib.tccg_gett_packed_noalloc.argtypes = [ctypes.c_int, vp, vp, vp, f64, f64, ctypes.POINTER(_CPlan), vp]
        lib.tccg_gett_packed_noalloc.restype = ctypes.c_int
        lib.tccg_gett_group.argtypes = [ctypes.c_int, ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_int, vp,
                                        ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_size_t, vp]
        lib.tccg_gett_group.restype = ctypes.c_int
        lib.tccg_packed_workspace_bytes.argtypes = [ctypes.c_int, ctypes.POINTER(_CPlan)]
        lib.tccg_packed_workspace_bytes.restype = ctypes.c_size_t
        lib.tccg_packed_tile_shape.argtypes = lib.tccg_tile_shape.argtypes
        lib.tccg_num_packed_tiles.argtypes = [ctypes.c_int]
        lib.tccg_vector_tile_shape.argtypes = lib.tccg_tile_shape.argtypes
        lib.tccg_num_vector_tiles.argtypes = [ctypes.c_int]
        lib.tccg_amax_f32.argtypes = [vp, ctypes.c_int64, vp, vp]
        from .transpose import TRANSPOSE_TILES, _CTransposePlan
        for name, T in (("tccg_transpose_f32", f32), ("tccg_transpose_f64", f64)):
            fn = getattr(lib, name)
            fn.argtypes = [vp, vp, T, T, ctypes.POINTER(_CTransposePlan), vp]
            fn.restype = ctypes.c_int
        lib.tccg_transpose_tile_shape.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        if lib.tccg_num_transpose_tiles() != len(TRANSPOSE_TILES):
            raise GettRuntimeError("transpose tile table mismatch")
        for t, shape in enumerate(TRANSPOSE_TILES):
            tx, ty = ctypes.c_int(), ctypes.c_int()
            lib.tccg_transpose_tile_shape(t, ctypes.byref(tx), ctypes.byref(ty))
            if (tx.value, ty.value) != shape:
                raise GettRuntimeError("transpose tile %d shape mismatch" % t)
        lib.tccg_gett_batch_f32.argtypes = [vp, vp, ctypes.c_int, ctypes.c_int, vp]
        lib.tccg_gett_batch_f32.restype = ctypes.c_int
        lib.tccg_batch_variant.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_int)] * 5
        lib.tccg_last_error.restype = ctypes.c_char_p
        lib.tccg_packed_sw_available.argtypes = [ctypes.c_int, ctypes.c_int]
        from .plan import sw_built
        for prec, tiles in PACKED_TILES.items():
            for t in range(len(tiles)):
                if bool(lib.tccg_packed_sw_available(prec, t)) != sw_built(prec, t):
                    raise GettRuntimeError("store-wave availability of packed tile %d/%d disagrees with the host"
                                           % (prec, t))
        # The Python tile tables must agree with the library's.
        for table, count, shape_of, kind in ((TILES, lib.tccg_num_tiles, lib.tccg_tile_shape, "tile"),
                                              (PACKED_TILES, lib.tccg_num_packed_tiles, lib.tccg_packed_tile_shape,
                                               "packed tile"),
                                              (VECTOR_TILES, lib.tccg_num_vector_tiles, lib.tccg_vector_tile_shape,
                                               "vector tile")):
            for prec, tiles in table.items():
                n = count(prec)
                if n != len(tiles):
                    raise GettRuntimeError("%s table mismatch for precision %d: library %d, host %d"
                                           % (kind, prec, n, len(tiles)))
                for t, shape in enumerate(tiles):
                    bm, bn, bk = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
                    shape_of(prec, t, ctypes.byref(bm), ctypes.byref(bn), ctypes.byref(bk))
                    if (bm.value, bn.value, bk.value) != shape:
                        raise GettRuntimeError("%s %d/%d shape mismatch" % (kind, prec, t))
        _lib = lib
        return lib


def last_error() -> str:
    return load_library().tccg_last_error().decode()


def _torch():
    import torch
    return torch


def check_tensor(name: str, t, dtype, numel: int, device) -> None:
    """A kernel argument must be a contiguous tensor of `dtype` with `numel`
    elements on `device` (the device the plan's tables live on): anything
    else would be read or written out of bounds or across devices."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.device.type != "cuda" or t.device.index != device.index:
        raise GettRuntimeError("%s is on %s; the plan's tables are on %s" % (name, t.device, device))
    if t.dtype != dtype:
        raise TypeError("%s has dtype %s, expected %s" % (name, t.dtype, dtype))
    if not t.is_contiguous():
        raise GettRuntimeError("%s must be contiguous (TCCG layout, leftmost index stride-one)" % name)
    if t.numel() != numel:
        raise GettRuntimeError("%s has %d elements, %d expected" % (name, t.numel(), numel))


def cuda_device(torch, device):
    """torch.device of `device` (default: the current one) with its index set."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise GettRuntimeError("the gfx950 GETT path runs on GPU devices, not %s" % dev)
    return dev if dev.index is not None else torch.device("cuda", torch.cuda.current_device())


def _check_cuda():
    torch = _torch()
    if not torch.cuda.is_available():
        raise GettRuntimeError("the gfx950 GETT path needs a ROCm GPU; none is visible")
    return torch


class DevicePlan:
    """A GettPlan whose tables live on the GPU."""

    def __init__(self, plan: GettPlan, device=None, workspace=None):
        """workspace: for packed plans, a device uint8 tensor of at least the
        plan's workspace bytes to use instead of a new one (callers that run
        several packed plans one after the other can share one)."""
        torch = _check_cuda()
        self.plan = plan
        dev = cuda_device(torch, device)
        self.device = dev

        def up(a: np.ndarray, dt):
            return torch.from_numpy(np.ascontiguousarray(a).view(dt)).to(dev)

        i32 = np.int32
        self.tables = [up(plan.offAm, i32), up(plan.offAk, i32), up(plan.offBn, i32), up(plan.offBk, i32),
                       up(plan.offCm, i32), up(plan.offCn, i32), up(plan.ordA, i32), up(plan.ordB, i32)]
        self.amax = torch.zeros(2, dtype=torch.float32, device=dev) if plan.prec == FP8 else None
        c = _CPlan()
        (c.offAm, c.offAk, c.offBn, c.offBk, c.offCm, c.offCn, c.ordA, c.ordB) = [t.data_ptr() for t in self.tables]
        c.Mp, c.Np, c.Kp = plan.Mp, plan.Np, plan.Kp
        c.tile, c.modeA, c.modeB, c.groupM = plan.tile, plan.modeA, plan.modeB, plan.groupM
        c.grid = plan.grid
        c.walk = plan.walk
        # the run epilogue is built for one tile (or one stream-k piece) at a
        # time; the store-wave kernel (walk 4) reads crun_m as its run length
        if plan.walk == 4:
            c.crun_m, c.crun_n = plan.crun[0], 0
        elif plan.walk == 5:
            c.crun_m, c.crun_n = plan.crun
        else:
            c.crun_m, c.crun_n = plan.crun if not plan.grid else (0, 0)
        c.cvec = plan.cvec
        c.ksplit = plan.ksplit
        c.amax = self.amax.data_ptr() if self.amax is not None else None
        self.cplan = c
        self.workspace = None
        if plan.packed:
            nbytes = plan.workspace_bytes()
            if load_library().tccg_packed_workspace_bytes(plan.prec, ctypes.byref(c)) != nbytes:
                raise GettRuntimeError("packed workspace size disagrees with the library")
            if workspace is not None:
                if workspace.dtype != torch.uint8 or workspace.numel() < nbytes or workspace.device != dev:
                    raise GettRuntimeError("shared workspace must be a uint8 tensor of >= %d bytes on %s"
                                           % (nbytes, dev))
                self.workspace = workspace
            else:
                self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)

    def _operands(self, A, B, C):
        """Tensors in the .tccg's (A, B, C) order -> kernel operands (opA, opB)."""
        pb = self.plan.problem
        torch = _torch()
        want = torch.float64 if self.plan.prec == F64 else torch.float32
        for name, t, T in (("A", A, pb.A), ("B", B, pb.B), ("C", C, pb.C)):
            check_tensor(name, t, want, T.size(), self.device)
        return (B, A) if pb.swapped else (A, B)

    def run(self, A, B, C, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        """C <- alpha * contract(A, B) + beta * C, asynchronously on `stream`
        (default: torch's current stream)."""
        torch = _torch()
        opA, opB = self._operands(A, B, C)
        with torch.cuda.device(self.device):
            self._launch(opA, opB, C, alpha, beta, stream)

    def _launch(self, opA, opB, C, alpha, beta, stream) -> None:
        torch = _torch()
        lib = load_library()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        sptr = ctypes.c_void_p(s.cuda_stream)
        prec = self.plan.prec
        if prec == FP8:
            rc = lib.tccg_amax_f32(opA.data_ptr(), opA.numel(), self.amax.data_ptr(), sptr)
            if rc == 0:
                rc = lib.tccg_amax_f32(opB.data_ptr(), opB.numel(), self.amax.data_ptr() + 4, sptr)
            if rc != 0:
                raise GettRuntimeError(last_error())
        if self.plan.packed:
            fn = getattr(lib, _PACKED_FN[prec])
            rc = fn(opA.data_ptr(), opB.data_ptr(), C.data_ptr(), alpha, beta, ctypes.byref(self.cplan),
                    self.workspace.data_ptr(), self.workspace.numel(), sptr)
        else:
            fn = {F32: lib.tccg_gett_f32, F64: lib.tccg_gett_f64, BF16X3: lib.tccg_gett_bf16x3,
                  FP8: lib.tccg_gett_fp8, BF16: lib.tccg_gett_bf16}[prec]
            rc = fn(opA.data_ptr(), opB.data_ptr(), C.data_ptr(), alpha, beta, ctypes.byref(self.cplan), sptr)
        if rc != 0:
            raise GettRuntimeError("GETT launch failed: %s" % last_error())


    def run_prepacked(self, C, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        """Packed plans only: the GETT kernel alone over the operands the last
        run() packed into this plan's workspace (A and B unchanged since)."""
        torch = _torch()
        if not self.plan.packed:
            raise GettRuntimeError("run_prepacked() needs a packed plan")
        check_tensor("C", C, torch.float64 if self.plan.prec == F64 else torch.float32, self.plan.problem.C.size(),
                     self.device)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        with torch.cuda.device(self.device):
            rc = load_library().tccg_gett_prepacked(self.plan.prec, C.data_ptr(), alpha, beta,
                                                    ctypes.byref(self.cplan), self.workspace.data_ptr(),
                                                    self.workspace.numel(), ctypes.c_void_p(s.cuda_stream))
        if rc != 0:
            raise GettRuntimeError("GETT launch failed: %s" % last_error())


class GeneratedGett:
    """A per-contraction library emitted by tccg_amd.codegen (exports gett,
    gett_stream, gett_info, gett_contraction, gett_last_error)."""

    INFO_FIELDS = ("M", "N", "K", "prec", "tile", "BM", "BN", "BK", "modeA", "modeB", "sizeA", "sizeB", "sizeC")

    def __init__(self, path):
        self.path = Path(path)
        self.lib = ctypes.CDLL(str(self.path))
        nf = len(self.INFO_FIELDS)
        buf = (ctypes.c_int64 * nf)()
        self.lib.gett_info.argtypes = [ctypes.c_void_p, ctypes.c_int]
        n = self.lib.gett_info(buf, nf)
        self.info = dict(zip(self.INFO_FIELDS, list(buf)[:n]))
        if n < nf:
            raise GettRuntimeError("%s reports %d plan fields, %d expected (regenerate it)" % (self.path, n, nf))
        T = ctypes.c_double if self.info["prec"] == F64 else ctypes.c_float
        self.lib.gett_stream.argtypes = [ctypes.c_void_p] * 3 + [T, T, ctypes.c_void_p]
        self.lib.gett_stream.restype = ctypes.c_int
        self.lib.gett.argtypes = [ctypes.c_void_p] * 3 + [T, T]
        self.lib.gett.restype = None
        self.lib.gett_contraction.restype = ctypes.c_char_p
        self.lib.gett_last_error.restype = ctypes.c_char_p

    def contraction(self) -> str:
        return self.lib.gett_contraction().decode()

    def __call__(self, A, B, C, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        torch = _check_cuda()
        # the library writes C and reads A and B with the extents fixed at
        # generation time: a tensor of another type, size or device would be
        # accessed out of bounds
        dev = cuda_device(torch, None)
        want = torch.float64 if self.info["prec"] == F64 else torch.float32
        for name, t in (("A", A), ("B", B), ("C", C)):
            check_tensor(name, t, want, self.info["size" + name], dev)
        s = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.gett_stream(A.data_ptr(), B.data_ptr(), C.data_ptr(), alpha, beta, ctypes.c_void_p(s.cuda_stream))
        if rc != 0:
            raise GettRuntimeError("generated gett failed: %s" % self.lib.gett_last_error().decode())
