"""Seeded random contractions for the parity sweeps (imported by tests; no
tests of its own): free indices of A only, of B only and contracted indices,
any of the three sets possibly empty as long as A, B and C keep an index,
every tensor's index order shuffled, extents 1-47 with every tensor under
2^18 elements."""
import numpy as np

from tccg_amd import Contraction

LIMIT = 1 << 18


def random_contraction(seed: int) -> Contraction:
    rng = np.random.default_rng(1000 + seed)
    while True:
        nm, nn, nk = (int(x) for x in rng.integers(0, 4, 3))
        if nm + nn == 0 or nm + nk == 0 or nn + nk == 0:
            continue
        labels = list("abcdefghij")
        rng.shuffle(labels)
        m, n, k = labels[:nm], labels[nm:nm + nn], labels[nm + nn:nm + nn + nk]
        sizes = {l: int(rng.integers(1, 48)) for l in m + n + k}
        prod = lambda ls: int(np.prod([sizes[l] for l in ls])) if ls else 1
        if max(prod(m + n), prod(m + k), prod(n + k)) > LIMIT:
            continue
        c_l, a_l, b_l = list(m + n), list(m + k), list(n + k)
        for ls in (c_l, a_l, b_l):
            rng.shuffle(ls)
        return Contraction.from_labels(c_l, a_l, b_l, sizes)
