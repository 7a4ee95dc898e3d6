"""Compilation of the gfx950 HIP sources into in-tree shared libraries.

The reference compiles its generated C++ with make and icpc/g++
(tccg/Makefile:1-51, driven from tccg/tccg.py:244-276); this module is the
gfx950 counterpart: hipcc --offload-arch=gfx950, one object per translation
unit built in parallel, linked into a shared library that the ctypes layer
loads.
"""
from __future__ import annotations

import concurrent.futures as _cf
import hashlib
import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_DIR = PKG_DIR / "_lib"
INCLUDE = PKG_DIR.parent / "include"
LIB_NAME = "libtccg_gfx950.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = "gfx950"
# Packed FP32 VALU forms are kept out of the generated code.
HIP_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

LIB_UNITS = ["gett_common.hip", "gett_f32.hip", "gett_f64.hip", "gett_bf16x3.hip", "gett_fp8.hip", "gett_bf16.hip",
             "gett_batch.hip", "gett_packed_f32.hip", "gett_packed_f64.hip", "gett_packed_bf16x3.hip",
             "gett_packed_bf16.hip", "gett_packed_fp8.hip", "transpose.hip", "gett_noalloc.hip",
             "gett_group_f32.hip", "gett_group_f64.hip", "gett_group_bf16x3.hip", "gett_group_bf16.hip"]
LIB_HEADERS = ["gett_core.hpp", "gett_launch.hpp", "gett_pack.hpp", "gett_packed.hpp", "gett_packed_sw.hpp",
               "gett_vector.hpp", "transpose.hpp", "gett_group.hpp"]


class CompilationError(RuntimeError):
    """hipcc failed; the message holds the command and compiler output."""


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
        raise CompilationError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def hipcc_available() -> bool:
    return Path(HIPCC).exists() or shutil.which("hipcc") is not None


def _sources_digest(files: list[Path]) -> str:
    h = hashlib.sha256()
    for f in files:
        h.update(f.name.encode())
        h.update(f.read_bytes())
    h.update(" ".join(HIP_FLAGS).encode())
    return h.hexdigest()[:16]


def library_path() -> Path:
    return LIB_DIR / LIB_NAME


def build_library(jobs: int | None = None, force: bool = True, verbose: bool = False) -> Path:
    """Compile the plan-driven GETT library; returns the path of the .so."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    deps = [CSRC / u for u in LIB_UNITS] + [CSRC / h for h in LIB_HEADERS] + [INCLUDE / "tccg_gfx950.h"]
    digest = _sources_digest(deps)
    stamp = LIB_DIR / (LIB_NAME + ".digest")
    out = library_path()
    if not force and out.exists() and stamp.exists() and stamp.read_text().strip() == digest:
        return out
    obj_dir = LIB_DIR / "obj"
    obj_dir.mkdir(exist_ok=True)
    jobs = jobs or min(len(LIB_UNITS), max(1, (os.cpu_count() or 2)))

    def compile_unit(unit: str) -> Path:
        obj = obj_dir / (Path(unit).stem + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I", str(INCLUDE), "-c", str(CSRC / unit), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with _cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_unit, LIB_UNITS))
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, out)
    stamp.write_text(digest + "\n")
    return out


PROBE_NAME = "libtccg_mfma_probe.so"


def probe_library_path() -> Path:
    return LIB_DIR / PROBE_NAME


def build_probe(verbose: bool = False) -> Path:
    """The MFMA-rate probe of the measurement harness (csrc/mfma_probe.hip), a
    library of its own: bench.py loads it, the GETT path never does."""
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    out = probe_library_path()
    tmp = out.with_suffix(".so.tmp")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-fPIC", str(CSRC / "mfma_probe.hip"), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, out)
    return out
