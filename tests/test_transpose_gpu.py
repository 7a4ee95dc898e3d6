"""Parity of the transposition kernels (tccg_transpose_f32 / _f64) with the
FP64 oracle, over every tile; each case's addressing is modelled on the CPU
by tests/test_transpose.py."""
import numpy as np
import pytest

from oracle.einsum_ref import rel_frobenius, transpose_ref
from tccg_amd.transpose import TRANSPOSE_TILES, DeviceTranspose, fuse_groups, plan_transpose

from .test_transpose import CASES, _rand

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ci", range(len(CASES)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_transpose_every_tile(ci, dtype):
    import torch
    i, o, sizes = CASES[ci]
    X = _rand(i, sizes, ci).astype(dtype)
    OUT = _rand(o, sizes, ci + 50).astype(dtype)
    dev = torch.device("cuda", 0)
    tX = torch.from_numpy(X).to(dev)
    tiles = [0] if fuse_groups(i, o)[0][0] == o[0] else range(1, len(TRANSPOSE_TILES))
    tol = 1e-6 if dtype == np.float32 else 1e-14
    for t in tiles:
        plan = plan_transpose(i, o, sizes, tile=t, elem_bytes=np.dtype(dtype).itemsize)
        print(plan.describe(), flush=True)
        dt = DeviceTranspose(plan, dev)
        tO = torch.from_numpy(OUT).to(dev)
        dt.run(tX, tO)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(tO.cpu().numpy(), transpose_ref(i, o, X).astype(dtype))   # a pure move
        tO = torch.from_numpy(OUT).to(dev)
        dt.run(tX, tO, 1.5, -0.25)
        torch.cuda.synchronize()
        assert rel_frobenius(tO.cpu().numpy(), transpose_ref(i, o, X, 1.5, -0.25, OUT)) < tol


def test_transpose_convenience_and_errors():
    import torch
    from tccg_amd.runtime import GettRuntimeError
    from tccg_amd.transpose import transpose
    X = torch.randn(7, 5, 3, device="cuda")           # TCCG a=3, b=5, c=7
    Y = transpose(X, "abc", "cab")
    assert tuple(Y.shape) == (5, 3, 7)
    np.testing.assert_array_equal(Y.cpu().numpy(), transpose_ref("abc", "cab", X.cpu().numpy()).astype(np.float32))
    dt = DeviceTranspose(plan_transpose("abc", "cab", dict(a=3, b=5, c=7)))
    with pytest.raises(GettRuntimeError):
        dt.run(X, torch.empty(10, device="cuda"))
