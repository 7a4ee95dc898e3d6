"""tccg_amd: MI355X (gfx950) GETT backend for TCCG tensor contractions.

Host side in Python (parser, index fusion, GETT planning), device side in
hand-written HIP (csrc/), reached through a C ABI (include/tccg_gfx950.h)
with ctypes.
"""
from .ir import Index, Tensor
from .frontend import TccgSyntaxError, fuse_indices, parse_tccg
from .plan import (BF16X3, F32, F64, FP8, GettPlan, GettProblem, UnsupportedContraction, build_plan,
                   candidate_plans, default_plan)
from .contraction import Contraction

__version__ = "0.1.0"

__all__ = [
    "Index", "Tensor", "TccgSyntaxError", "fuse_indices", "parse_tccg", "GettPlan", "GettProblem",
    "UnsupportedContraction", "build_plan", "candidate_plans", "default_plan", "Contraction",
    "F32", "F64", "BF16X3", "FP8",
]
