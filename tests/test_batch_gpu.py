"""Batched persistent GETT: parity per problem and HIP-graph replays."""
import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd.batch import GettBatch, autotune_patterns, random_batch

pytestmark = pytest.mark.gpu


def make_tensors(cs, seed):
    import torch
    rng = np.random.default_rng(seed)
    As, Bs, Cs, hostA, hostB = [], [], [], [], []
    for c in cs:
        sa, sb, sc = c.shapes()
        A = rng.uniform(-1, 1, sa).astype(np.float32)
        B = rng.uniform(-1, 1, sb).astype(np.float32)
        hostA.append(A)
        hostB.append(B)
        As.append(torch.from_numpy(A).cuda())
        Bs.append(torch.from_numpy(B).cuda())
        Cs.append(torch.zeros(sc, dtype=torch.float32, device="cuda"))
    return As, Bs, Cs, hostA, hostB


def lab(T):
    return [i.label for i in T.indices]


@pytest.mark.parametrize("shapes", [None, "tuned"])
def test_batch_parity(shapes):
    import torch
    cs = random_batch(48, seed=7, hi=48)
    As, Bs, Cs, hA, hB = make_tensors(cs, 1)
    sh = autotune_patterns(cs, As, Bs, Cs, reps=1) if shapes == "tuned" else None
    b = GettBatch(cs, sh).bind(As, Bs, Cs)
    b.launch()
    torch.cuda.synchronize()
    for i, c in enumerate(cs):
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), hA[i], hB[i])
        assert rel_frobenius(Cs[i].cpu().numpy(), ref) < 1e-5, (i, str(c))


def test_batch_alpha_beta_and_ragged():
    import torch
    from tccg_amd import Contraction
    cs = [Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], dict(a=5, b=7, k=3)),
          Contraction.from_labels(["a", "b", "c"], ["c", "k", "a"], ["k", "b"], dict(a=13, b=6, c=9, k=17)),
          Contraction.from_labels(["b", "a"], ["a", "k"], ["k", "b"], dict(a=40, b=33, k=65))]
    As, Bs, Cs, hA, hB = make_tensors(cs, 2)
    C0 = [torch.randn_like(C) for C in Cs]
    for C, c0 in zip(Cs, C0):
        C.copy_(c0)
    b = GettBatch(cs, alphas=[2.0, -1.0, 0.5], betas=[0.5, 1.0, 0.0]).bind(As, Bs, Cs)
    b.launch()
    torch.cuda.synchronize()
    for i, c in enumerate(cs):
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), hA[i], hB[i], C0[i].cpu().numpy(),
                           [2.0, -1.0, 0.5][i], [0.5, 1.0, 0.0][i])
        assert rel_frobenius(Cs[i].cpu().numpy(), ref) < 1e-5


def test_batch_graph_replays():
    import torch
    cs = random_batch(32, seed=9, hi=40)
    As, Bs, Cs, hA, hB = make_tensors(cs, 3)
    b = GettBatch(cs).bind(As, Bs, Cs)
    b.launch()
    torch.cuda.synchronize()
    eager = [C.clone() for C in Cs]
    g = b.capture()
    for r in range(3):
        for C in Cs:
            C.zero_()                        # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for C, E in zip(Cs, eager):
            assert torch.equal(C, E), "replay %d differs from the eager launch" % r


def test_batch_with_packed_problems():
    """Problems with enough tiles run on the packed path (shared workspace)
    ahead of the persistent launch; every problem still matches the oracle,
    eagerly and in graph replays."""
    import torch
    cs = random_batch(40, seed=11, hi=64)
    As, Bs, Cs, hA, hB = make_tensors(cs, 4)
    b = GettBatch(cs, packed_min_tiles=16)
    assert 0 < len(b.large) < len(cs)
    b.bind(As, Bs, Cs)
    b.launch()
    torch.cuda.synchronize()
    for i, c in enumerate(cs):
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), hA[i], hB[i])
        assert rel_frobenius(Cs[i].cpu().numpy(), ref) < 1e-5, (i, i in b.large, str(c))
    eager = [C.clone() for C in Cs]
    g = b.capture()
    for r in range(3):
        for C in Cs:
            C.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for C, E in zip(Cs, eager):
            assert torch.equal(C, E), "replay %d differs from the eager launch" % r
