"""The allocation-free reduction (nrs_linalg_reduce_noalloc) against
nrs_linalg_reduce, through the C ABI: every reduction, one and three fields,
two calls in a row with the second on a smaller input (the fixed scratch keeps
the first call's partials, which the second must not read), bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = ["sum", "max", "min", "norm2", "norm1", "amax", "weighted_norm2", "weighted_inner_prod",
       "inner_prod"]


def _inputs(torch, N, nfields, seed):
    rng = np.random.default_rng(seed)
    off = N if nfields == 1 else ((N + 127) // 128) * 128
    x = torch.from_numpy(rng.uniform(-1, 1, nfields * off)).cuda()
    y = torch.from_numpy(rng.uniform(-1, 1, nfields * off)).cuda()
    w = torch.from_numpy(rng.uniform(0.1, 1, N)).cuda()
    return x, y, w, off


@pytest.mark.parametrize("nfields", [1, 3])
@pytest.mark.parametrize("sizes", [(3_000_001, 70_001), (65_537, 5), (16 * 256, 1)])
def test_noalloc_reduce_matches_reduce(gpu, torch_mod, nfields, sizes):
    torch = torch_mod
    for op in OPS:
        for i, N in enumerate(sizes):   # a large call, then a smaller one on the same scratch
            x, y, w, off = _inputs(torch, N, nfields, 100 * i + nfields)
            ref = gpu.linalg_reduce(op, x, y, w, N=N, nfields=nfields, field_offset=off)
            got = gpu.linalg_reduce(op, x, y, w, N=N, nfields=nfields, field_offset=off,
                                    noalloc=True)
            assert got == ref or (np.isnan(got) and np.isnan(ref)), (op, N, nfields, got, ref)


def test_noalloc_reduce_empty_and_errors(gpu, torch_mod):
    torch = torch_mod
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert gpu.linalg_reduce("sum", x, N=0, noalloc=True) == 0.0
    assert gpu.linalg_reduce("max", x, N=0, noalloc=True) == -np.inf
    with pytest.raises(gpu._lib.NrsError):
        gpu.linalg_reduce("inner_prod", x, None, N=4, noalloc=True)   # null y is refused
