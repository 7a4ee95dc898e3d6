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


def _checksum(c, A, B):
    """sum over C of A.B = sum_k (sum over A's free indices of A)(sum over B's
    free indices of B), in FP64: a size-independent property of a contraction."""
    ka = [i.label for i in c.A0.indices]
    kb = [i.label for i in c.B0.indices]
    k = [l for l in ka if l in kb]
    ra = np.asarray(A, np.float64).sum(axis=tuple(A.ndim - 1 - ka.index(l) for l in ka if l not in k))
    rb = np.asarray(B, np.float64).sum(axis=tuple(B.ndim - 1 - kb.index(l) for l in kb if l not in k))
    # remaining axes are the k indices in reversed A / B order
    ka_rev = [l for l in reversed(ka) if l in k]
    kb_rev = [l for l in reversed(kb) if l in k]
    rb = np.transpose(rb, [kb_rev.index(l) for l in ka_rev])
    return float((ra * rb).sum()), float((np.abs(ra) * np.abs(rb)).sum())


@pytest.mark.parametrize("mode", ["persistent", "grouped"])
def test_full_batch_every_problem(mode):
    """The benchmark's batch: 512 contractions of the 4 patterns, every extent
    uniform over multiples of 8 in [8, 96] (seed 0), either with the large
    problems on the packed path ahead of the persistent launch, or all of
    them in one grouped packed launch (pack + GETT).  Every problem: against
    the FP64 einsum oracle on the CPU (problems up to 2 GFLOP) or FP64
    torch.einsum on the GPU (the rest, an independent library path), and its
    element sum against the FP64 checksum."""
    import torch
    from tccg_amd.plan import F32, PACKED_TILES
    cs = random_batch(512, seed=0)
    As, Bs, Cs, hA, hB = make_tensors(cs, 0)
    if mode == "grouped":
        b = GettBatch(cs, group_tile=PACKED_TILES[F32].index((128, 128, 16))).bind(As, Bs, Cs)
        assert b.launches() == 2
        print("full batch grouped: %d problems, %d tiles" % (len(cs), b._bound["group"]._bound["nwork"]), flush=True)
    else:
        b = GettBatch(cs, packed_min_tiles=512).bind(As, Bs, Cs)
        print("full batch: %d problems, %d on the packed path, %d tiles" % (len(cs), len(b.large),
                                                                            b._bound["nwork"]), flush=True)
    g = b.capture()
    g.replay()
    torch.cuda.synchronize()
    worst = 0.0
    for i, c in enumerate(cs):
        C = Cs[i]
        if c.flops() <= 2e9:
            ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), hA[i], hB[i])
            err = rel_frobenius(C.cpu().numpy(), ref)
        else:
            import string
            letters = dict(zip(sorted({l for T in (c.C0, c.A0, c.B0) for l in lab(T)}), string.ascii_letters))
            rev = lambda T: "".join(letters[l] for l in reversed(lab(T)))
            ref = torch.einsum("%s,%s->%s" % (rev(c.A0), rev(c.B0), rev(c.C0)), As[i].double(), Bs[i].double())
            err = float((C.double() - ref).norm() / ref.norm())
            del ref
        worst = max(worst, err)
        assert err < 1e-5, (i, mode, str(c), err)
        s, scale = _checksum(c, hA[i], hB[i])
        got = float(C.double().sum())
        assert abs(got - s) <= 1e-5 * scale, (i, str(c), got, s)
    print("full batch: worst rel. Frobenius error %.3e" % worst, flush=True)
