"""Writes the golden fixtures of the GETT path: seeded inputs and FP64 outputs.

The reference stores no golden vectors (its only check is the run-time
comparison with a naive loop nest, tccg/tccg.py:1010-1050), so the fixtures
are produced here from numpy.einsum in FP64 and cross-checked, before
writing, against a brute-force loop nest in plain Python; each fixture is a
small .npz (inputs A, B, C_in; scalars; expected C_out in float64).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import itertools
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.einsum_ref import contract_ref  # noqa: E402

CASES = {
    # name: (C, A, B, sizes, alpha, beta, seed)
    "example_small": (["a", "b", "i", "j"], ["i", "m", "a"], ["m", "j", "b"], dict(a=6, b=5, i=4, j=3, m=7), 1.0, 0.0, 0),
    "cpu48_small": (["a", "b", "c", "d"], ["a", "k", "c"], ["k", "d", "b"], dict(a=8, b=5, c=6, d=4, k=9), 1.0, 0.0, 1),
    "aebf_dfce_small": (["a", "b", "c", "d"], ["a", "e", "b", "f"], ["d", "f", "c", "e"], dict(a=6, b=5, c=4, d=7, e=3, f=5), 1.0, 0.0, 2),
    "akb_kc_small": (["a", "b", "c"], ["a", "k", "b"], ["k", "c"], dict(a=12, b=5, c=7, k=8), 1.0, 0.0, 3),
    "dgfa_gceb_small": (["a", "b", "c", "d", "e", "f"], ["d", "g", "f", "a"], ["g", "c", "e", "b"], dict(a=3, b=4, c=2, d=5, e=3, f=2, g=6), 1.0, 0.0, 4),
    "aklb_lckd_small": (["a", "b", "c", "d"], ["a", "k", "l", "b"], ["l", "c", "k", "d"], dict(a=4, b=5, c=3, d=6, k=2, l=3), 1.0, 0.0, 5),
    "transC_alpha_beta": (["i1", "j", "i2"], ["i1", "k", "i2"], ["j", "k"], dict(i1=8, i2=5, j=6, k=4), 2.4, 1.1, 6),
}


def brute(c, a, b, sizes, A, B, C, alpha, beta):
    out = np.zeros(C.shape, dtype=np.float64)
    k = [l for l in a if l in b and l not in c]
    for fidx in itertools.product(*[range(sizes[l]) for l in c]):
        env = dict(zip(c, fidx))
        s = 0.0
        for kidx in itertools.product(*[range(sizes[l]) for l in k]):
            env.update(zip(k, kidx))
            s += float(A[tuple(env[l] for l in reversed(a))]) * float(B[tuple(env[l] for l in reversed(b))])
        pos = tuple(env[l] for l in reversed(c))
        out[pos] = alpha * s + beta * float(C[pos])
    return out


def make(name):
    c, a, b, sizes, alpha, beta, seed = CASES[name]
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, [sizes[l] for l in reversed(a)]).astype(np.float32)
    B = rng.uniform(-1, 1, [sizes[l] for l in reversed(b)]).astype(np.float32)
    C = rng.uniform(-1, 1, [sizes[l] for l in reversed(c)]).astype(np.float32)
    ref = contract_ref(c, a, b, A, B, C, alpha, beta)
    bf = brute(c, a, b, sizes, A, B, C, alpha, beta)
    assert np.max(np.abs(ref - bf)) < 1e-12, name
    return dict(A=A, B=B, C_in=C, C_out=ref, alpha=np.float64(alpha), beta=np.float64(beta),
                c=np.array(c), a=np.array(a), b=np.array(b),
                size_labels=np.array(list(sizes)), size_values=np.array(list(sizes.values()), dtype=np.int64))


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **make(name))
        print("wrote", name)
