"""Named .tccg workloads: the reference's example and the measured configurations.

Sizes follow the benchmark definitions of this backend (one GETT launch per
workload); `scaled(name, f)` shrinks every extent for quick parity runs.
"""
from __future__ import annotations

import re

WORKLOADS = {
    # reference example.tccg:1-6
    "example": """C[a,b,i,j] = A[i,m,a] * B[m,j,b]
a = 24
b = 24
i = 24
j = 24
m = 24
""",
    # CPU configuration: C[a,b,c,d] = A[a,k,c] * B[k,d,b], 48/64
    "cpu48": """C[a,b,c,d] = A[a,k,c] * B[k,d,b]
a = 48
b = 48
c = 48
d = 48
k = 64
""",
    # compute-bound: the abcd-aebf-dfce pattern (benchmark/benchmark.py:55 of the reference) at 72/48
    "abcd-aebf-dfce": """C[a,b,c,d] = A[a,e,b,f] * B[d,f,c,e]
a = 72
b = 72
c = 72
d = 72
e = 48
f = 48
""",
    # bandwidth-bound: C[a,b,c] = A[a,k,b] * B[k,c], 384/8 (cf. benchmark/transC.py:5-9 of the reference)
    "abc-akb-kc": """C[a,b,c] = A[a,k,b] * B[k,c]
a = 384
b = 384
c = 384
k = 8
""",
    # precision study: 6-D CCSD(T)-like contraction, free extents 24, g = 96
    "abcdef-dgfa-gceb": """C[a,b,c,d,e,f] = A[d,g,f,a] * B[g,c,e,b]
a = 24
b = 24
c = 24
d = 24
e = 24
f = 24
g = 96
""",
    # differentiation study: C[a,b,c,d] = A[a,k,l,b] * B[l,c,k,d], 72/32
    "abcd-aklb-lckd": """C[a,b,c,d] = A[a,k,l,b] * B[l,c,k,d]
a = 72
b = 72
c = 72
d = 72
k = 32
l = 32
""",
}

# index patterns of the batched study: (C, A, B) label lists
BATCH_PATTERNS = [
    (["a", "b", "c"], ["a", "k", "b"], ["k", "c"]),
    (["a", "b", "c", "d"], ["a", "k", "l", "b"], ["l", "c", "k", "d"]),
    (["a", "b"], ["k", "a"], ["b", "k"]),
    (["a", "b", "c", "d"], ["d", "g", "a"], ["g", "c", "b"]),
]


def get(name: str) -> str:
    return WORKLOADS[name]


def scaled(name: str, factor: float, minimum: int = 2) -> str:
    """The workload with every extent multiplied by `factor` (rounded, >= minimum)."""
    def sub(m):
        v = max(minimum, int(round(int(m.group(2)) * factor)))
        return "%s = %d" % (m.group(1), v)
    return re.sub(r"^([a-z][a-z0-9]*)\s*=\s*([0-9]+)", sub, WORKLOADS[name], flags=re.M)
