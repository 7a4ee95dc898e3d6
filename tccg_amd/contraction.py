"""User-facing contraction object of the gfx950 backend.

    c = Contraction.from_tccg(open("example.tccg").read())
    c(A, B, C)                 # C <- alpha * A.B + beta * C on the GPU

A, B and C are contiguous torch tensors on the GPU holding the TCCG
(column-major, leftmost index stride-one) layout, i.e. torch shape
reversed(indices).  The alpha/beta of the .tccg are the defaults; both can be
overridden per call, as the generated gett(A, B, C, alpha, beta) of the
reference allows (tccg/gett.py:342-344).
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

from .frontend import contraction_str, fuse_indices, parse_tccg
from .ir import Index, Tensor
from .plan import F32, F64, PREC_BY_NAME, PREC_NAMES, GettPlan, GettProblem, build_plan, candidate_plans, default_plan


def _prec_id(prec) -> int:
    if isinstance(prec, str):
        key = {"s": "f32", "d": "f64", "float": "f32", "double": "f64"}.get(prec, prec)
        return PREC_BY_NAME[key]
    return int(prec)


def tensors_from_spec(c: Sequence[str], a: Sequence[str], b: Sequence[str], sizes: Dict[str, int]):
    mk = lambda lab, idx: Tensor(lab, [Index(l, sizes[l]) for l in idx])
    return mk("C", c), mk("A", a), mk("B", b)


class Contraction:
    """One contraction C = alpha * A.B + beta * C, after index fusion."""

    def __init__(self, C: Tensor, A: Tensor, B: Tensor, alpha: float = 1.0, beta: float = 0.0,
                 fuse: bool = True):
        import copy
        from .algebra import get_loop_indices
        from .plan import UnsupportedContraction
        # Batch indices are rejected before fusion: fusing a run that contains
        # one would leave it in a single tensor (the reference fuses first,
        # tccg/tccg.py:104, and only then checks, tccg/gett.py:42-45).
        for T in (A, B, C):
            for i in T.indices:
                if i.size < 1:
                    raise UnsupportedContraction("index %s of %s has extent %d: extents must be positive"
                                                 % (i.label, T.label, i.size))
            # Kernels add the uint32 per-set offsets of an element in 32 bits:
            # every element offset, hence every such sum, must stay below the
            # INVALID sentinel 2^32 - 1.
            if T.size() >= 0xFFFFFFFF:
                raise UnsupportedContraction("tensor %s has %d elements; at most 2^32 - 2 are addressable"
                                             % (T.label, T.size()))
        loop = get_loop_indices(A, B, C)
        if loop:
            raise UnsupportedContraction("loop indices are not yet supported by GETT: %s"
                                         % ",".join(i.label for i in loop))
        # the caller-visible (unfused) tensors define the expected element counts
        self.C0, self.A0, self.B0 = copy.deepcopy(C), copy.deepcopy(A), copy.deepcopy(B)
        self.C, self.A, self.B = copy.deepcopy(C), copy.deepcopy(A), copy.deepcopy(B)
        if fuse:
            fuse_indices(self.A, self.B, self.C)
        self.alpha, self.beta = float(alpha), float(beta)
        self.problem = GettProblem(self.C, self.A, self.B)
        self._device_plans: Dict[Tuple, object] = {}
        self._tuned: Dict[int, GettPlan] = {}

    @classmethod
    def from_tccg(cls, text: str, fuse: bool = True) -> "Contraction":
        C, A, B, alpha, beta = parse_tccg(text)
        return cls(C, A, B, alpha, beta, fuse=fuse)

    @classmethod
    def from_labels(cls, c: Sequence[str], a: Sequence[str], b: Sequence[str], sizes: Dict[str, int],
                    alpha: float = 1.0, beta: float = 0.0, fuse: bool = True) -> "Contraction":
        C, A, B = tensors_from_spec(c, a, b, sizes)
        return cls(C, A, B, alpha, beta, fuse=fuse)

    def __str__(self) -> str:
        return contraction_str(self.C, self.A, self.B, self.alpha, self.beta)

    def flops(self, beta: Optional[float] = None) -> float:
        return self.problem.flops(self.beta if beta is None else beta)

    def shapes(self):
        """torch shapes of (A, B, C) in the caller's (unfused) layout."""
        return self.A0.torch_shape(), self.B0.torch_shape(), self.C0.torch_shape()

    def plan(self, prec="f32", tile: Optional[int] = None, m_blocked: Optional[bool] = None) -> GettPlan:
        """The tuned plan (after tune()), else the model's best; an explicit
        tile / blocking builds that plan of the gather kernel."""
        p = _prec_id(prec)
        if tile is None and m_blocked is None:
            return self._tuned.get(p) or default_plan(self.problem, p)
        if tile is None:
            # `tile` indexes the gather kernel's table: the model's best plan of that kernel
            tile = next(q.tile for q in candidate_plans(self.problem, p) if not q.packed and q.walk != 3)
        return build_plan(self.problem, p, tile, m_blocked=m_blocked)

    def tune(self, prec="f32", max_implementations: int = 16, inputs=None):
        """Time the best-ranked candidate plans on the GPU (tccg_amd.autotune,
        the reference's measure-and-select step, tccg/tccg.py:188-430) and make
        the fastest this contraction's plan for `prec`; returns its measurement."""
        from .autotune import autotune
        p = _prec_id(prec)
        best = autotune(self, p, max_implementations=max_implementations, inputs=inputs)[0]
        self._tuned[p] = best.plan
        self._device_plans.pop((p, None, None), None)
        return best

    def device_plan(self, prec="f32", tile: Optional[int] = None, m_blocked: Optional[bool] = None,
                    plan: Optional[GettPlan] = None):
        from .runtime import DevicePlan
        if plan is not None:
            key = ("explicit", id(plan))
        else:
            key = (_prec_id(prec), tile, m_blocked)
        dp = self._device_plans.get(key)
        if dp is None:
            dp = DevicePlan(plan if plan is not None else self.plan(prec, tile, m_blocked))
            self._device_plans[key] = dp
        return dp

    def __call__(self, A, B, C, alpha: Optional[float] = None, beta: Optional[float] = None, prec="f32",
                 tile: Optional[int] = None, stream=None):
        dp = self.device_plan(prec, tile)
        dp.run(A, B, C, self.alpha if alpha is None else alpha, self.beta if beta is None else beta,
               stream=stream)
        return C
