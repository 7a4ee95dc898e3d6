"""The grouped packed GETT (tccg_gett_group: one pack launch and one GETT
launch over many contractions) against the FP64 einsum oracle: batches of
uneven contractions in every precision it runs and several tiles, with
alpha/beta and ragged extents, HIP-graph replays, and the host checks that
reject a call before any launch."""
import ctypes

import numpy as np
import pytest

from oracle.einsum_ref import contract_ref, rel_frobenius
from tccg_amd import Contraction
from tccg_amd.batch import random_batch
from tccg_amd.group import GettGroup, group_for_plans
from tccg_amd.plan import BF16, BF16X3, F32, F64, FP8, PACKED_TILES, build_packed_plan

pytestmark = pytest.mark.gpu

TOL = {F32: 1e-5, F64: 1e-13, BF16X3: 1e-5, BF16: 1e-2}


def lab(T):
    return [i.label for i in T.indices]


def _tensors(cs, seed, prec, with_c):
    import torch
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec == F64 else np.float32
    host, dev = [], []
    for c in cs:
        sa, sb, sc = c.shapes()
        A, B = rng.uniform(-1, 1, sa).astype(dt), rng.uniform(-1, 1, sb).astype(dt)
        C = rng.uniform(-1, 1, sc).astype(dt) if with_c else np.zeros(sc, dt)
        host.append((A, B, C))
        dev.append(tuple(torch.from_numpy(x).cuda() for x in (A, B, C)))
    return host, dev


def _run_group(cs, prec, tile, alphas, betas, seed=0):
    import torch
    host, dev = _tensors(cs, seed, prec, with_c=any(b != 0 for b in betas))
    plans = [build_packed_plan(c.problem, prec, tile) for c in cs]
    g = group_for_plans(plans, [d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev], alphas, betas)
    g.bind(torch.device("cuda", 0))
    g.launch()
    torch.cuda.synchronize()
    return g, host, dev


@pytest.mark.parametrize("prec,tile", [(F32, t) for t in range(len(PACKED_TILES[F32]))] +
                         [(F64, 0), (F64, 1), (BF16X3, 0), (BF16X3, 5), (BF16, 0)])
def test_group_batch(prec, tile):
    cs = random_batch(24, seed=5 + tile, hi=64)
    alphas = [1.0 + 0.25 * (i % 3) for i in range(len(cs))]
    betas = [0.0 if i % 2 else -0.5 for i in range(len(cs))]
    print("group", prec, PACKED_TILES[prec][tile], flush=True)
    g, host, dev = _run_group(cs, prec, tile, alphas, betas, seed=tile)
    for i, c in enumerate(cs):
        A, B, C = host[i]
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, alphas[i], betas[i])
        err = rel_frobenius(dev[i][2].cpu().numpy(), ref)
        assert err < TOL[prec], (i, str(c), prec, tile, err)


def test_group_ragged_and_swapped():
    cs = [Contraction.from_labels(["a", "b"], ["a", "k"], ["k", "b"], dict(a=5, b=7, k=3)),
          Contraction.from_labels(["a", "b", "c"], ["c", "k", "a"], ["k", "b"], dict(a=13, b=6, c=9, k=17)),
          Contraction.from_labels(["b", "a"], ["a", "k"], ["k", "b"], dict(a=40, b=33, k=65)),
          Contraction.from_labels(["a", "b", "c", "d"], ["d", "k", "a", "l"], ["l", "c", "k", "b"],
                                  dict(a=5, b=7, c=3, d=11, k=9, l=4))]
    tile = PACKED_TILES[F32].index((128, 128, 16))
    g, host, dev = _run_group(cs, F32, tile, [2.0, -1.0, 0.5, 1.0], [0.5, 1.0, 0.0, 2.0], seed=2)
    for i, c in enumerate(cs):
        A, B, C = host[i]
        ref = contract_ref(lab(c.C0), lab(c.A0), lab(c.B0), A, B, C, [2.0, -1.0, 0.5, 1.0][i],
                           [0.5, 1.0, 0.0, 2.0][i])
        assert rel_frobenius(dev[i][2].cpu().numpy(), ref) < 1e-5, str(c)


def test_group_graph_replays():
    import torch
    from tccg_amd.batch import GettBatch
    cs = random_batch(32, seed=9, hi=48)
    host, dev = _tensors(cs, 3, F32, with_c=False)
    b = GettBatch(cs, group_tile=PACKED_TILES[F32].index((128, 128, 16)))
    b.bind([d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev])
    b.launch()
    torch.cuda.synchronize()
    eager = [d[2].clone() for d in dev]
    graph = b.capture()
    for r in range(3):
        for d in dev:
            d[2].zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for d, E in zip(dev, eager):
            assert torch.equal(d[2], E), "replay %d differs from the eager launch" % r


def test_group_rejects_before_launch():
    import torch
    from tccg_amd.runtime import load_library
    lib = load_library()
    ws = torch.empty(1 << 20, dtype=torch.uint8, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    vp = ctypes.c_void_p
    fake = torch.zeros(64, dtype=torch.int32, device="cuda")
    args = lambda prec, tile, wsp, wsb: (prec, tile, vp(fake.data_ptr()), 1, vp(fake.data_ptr()), 1,
                                          vp(fake.data_ptr()), 1, vp(fake.data_ptr()), 1, wsp, wsb, s)
    assert lib.tccg_gett_group(*args(FP8, 0, vp(ws.data_ptr()), ws.numel())) != 0
    assert b"FP32, FP64" in lib.tccg_last_error()
    assert lib.tccg_gett_group(*args(F32, 99, vp(ws.data_ptr()), ws.numel())) != 0
    assert lib.tccg_gett_group(*args(F32, 0, vp(ws.data_ptr()), 1024)) != 0      # below the tail pad
    assert lib.tccg_gett_group(*args(F32, 0, vp(ws.data_ptr() + 4), ws.numel() - 4)) != 0   # misaligned
    torch.cuda.synchronize()
