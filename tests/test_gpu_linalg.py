"""The device reductions and the PCG residual update against host loops.

Restates the reference's ciTestLinAlg (examples/ethier/ci.inc:729-790): the
linAlg_t reductions sum, max, min, norm2, norm1, amax on uniform(0,1) vectors
of N = 1, BLOCKSIZE/16 ... 8 BLOCKSIZE (BLOCKSIZE = 256, ci.inc:842-850) and
weightedNorm2Many for 1..10 fields with the field offset padded to 1024 bytes,
each within 50 eps relative of a sequential host loop; extended here to 16
BLOCKSIZE, odd sizes, signed data, the inner products and the size of the
operator benchmark.  ellipticBlockUpdatePCG is checked against the golden
vectors recorded from the reference's compiled serial kernel
(tests/golden/updatepcg.npz) and against a host loop with three fields."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
TOL = 50 * EPS
BLOCKSIZE = 256
SIZES = [1, BLOCKSIZE // 16, BLOCKSIZE // 8, BLOCKSIZE // 4, BLOCKSIZE // 2, BLOCKSIZE,
         2 * BLOCKSIZE, 4 * BLOCKSIZE, 8 * BLOCKSIZE, 16 * BLOCKSIZE, 3, 1000, 65537]


def host(op, x, y=None, w=None, nfields=1, off=None):
    """Sequential loops as ci.inc:648-727 writes them (math.fsum where a loop
    over millions of entries would be slow; it is the exactly rounded sum)."""
    if op == "sum":
        return math.fsum(x.tolist())
    if op == "max":
        m = -9e30
        for v in x.tolist():
            m = v if v > m else m
        return m
    if op == "min":
        m = 9e30
        for v in x.tolist():
            m = v if v < m else m
        return m
    if op == "norm2":
        return math.sqrt(math.fsum((x * x).tolist()))
    if op == "norm1":
        return math.fsum(np.abs(x).tolist())
    if op == "amax":
        return float(np.max(np.abs(x)))
    if op == "inner_prod":
        return math.fsum((x * y).tolist())
    raise ValueError(op)


def field_offset(N):
    page = 1024 // 8
    return N if N % page == 0 else (N // page + 1) * page


@pytest.mark.parametrize("N", SIZES)
def test_reductions_vs_host_loops(gpu, torch_mod, N):
    torch = torch_mod
    rng = np.random.default_rng(N)
    for data in (rng.uniform(0.0, 1.0, N), rng.uniform(-1.0, 1.0, N) + 0.25):
        y = rng.uniform(0.0, 1.0, N)
        xd, yd = torch.from_numpy(data).cuda(), torch.from_numpy(y).cuda()
        for op in ("sum", "max", "min", "norm2", "norm1", "amax", "inner_prod"):
            ref = host(op, data, y)
            got = gpu.linalg_reduce(op, xd, yd)
            # terms of both signs: relative to the sum of magnitudes
            scale = host("norm1", data * y) if op == "inner_prod" else (
                host("norm1", data) if op == "sum" else abs(ref))
            err = abs(got - ref) / scale
            assert err <= TOL, (op, N, got, ref, err)
            if op in ("max", "min", "amax"):
                assert got == ref


@pytest.mark.parametrize("N", SIZES)
def test_weighted_many_vs_host_loops(gpu, torch_mod, N):
    """weightedNorm2Many / weightedInnerProdMany, Nfields = 1..10 (ci.inc:765-787)."""
    torch = torch_mod
    off = field_offset(N)
    rng = np.random.default_rng(1000 + N)
    for nf in range(1, 11):
        x = rng.uniform(0.0, 1.0, nf * off)
        y = rng.uniform(-1.0, 1.0, nf * off)
        w = rng.uniform(0.0, 1.0, N)
        xs = x.reshape(nf, off)[:, :N]
        ys = y.reshape(nf, off)[:, :N]
        ref2 = math.sqrt(math.fsum((w[None] * xs * xs).ravel().tolist()))
        refd = math.fsum((w[None] * xs * ys).ravel().tolist())
        xd, yd, wd = (torch.from_numpy(a).cuda() for a in (x, y, w))
        got2 = gpu.linalg_reduce("weighted_norm2", xd, w=wd, N=N, nfields=nf, field_offset=off)
        gotd = gpu.linalg_reduce("weighted_inner_prod", xd, yd, w=wd, N=N, nfields=nf, field_offset=off)
        assert abs(got2 - ref2) <= TOL * ref2, (nf, N, got2, ref2)
        # the inner product has terms of both signs: bound by the sum of magnitudes
        mag = math.fsum(np.abs(w[None] * xs * ys).ravel().tolist())
        assert abs(gotd - refd) <= TOL * mag, (nf, N, gotd, refd)


def test_reductions_large_and_deterministic(gpu, torch_mod):
    """The operator benchmark's vector length (32^3 elements of order 7): the
    same bits on every call, within 50 eps of the exactly rounded result."""
    torch = torch_mod
    n = 32 ** 3 * 512
    x = torch.rand(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    w = torch.rand(n, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(4))
    xh, wh = x.cpu().numpy(), w.cpu().numpy()
    ref = math.sqrt(math.fsum((wh * xh * xh).tolist()))
    a = gpu.linalg_reduce("weighted_norm2", x, w=w)
    b = gpu.linalg_reduce("weighted_norm2", x, w=w)
    assert a == b
    assert abs(a - ref) <= TOL * ref
    assert gpu.linalg_reduce("sum", x) == gpu.linalg_reduce("sum", x)


def test_empty_reductions(gpu, torch_mod):
    torch = torch_mod
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    assert gpu.linalg_reduce("sum", x, N=0) == 0.0
    assert gpu.linalg_reduce("max", x, N=0) == -math.inf
    assert gpu.linalg_reduce("min", x, N=0) == math.inf
    assert gpu.linalg_reduce("amax", x, N=0) == 0.0


def test_block_update_pcg_golden(gpu, torch_mod, golden):
    """Against the reference's compiled ellipticBlockUpdatePCG (one field)."""
    torch = torch_mod
    g = golden("updatepcg")
    r = torch.from_numpy(g["updatepcg_r"].copy()).cuda()
    w = torch.from_numpy(g["updatepcg_w"]).cuda()
    Ap = torch.from_numpy(g["updatepcg_Ap"]).cuda()
    alpha = float(g["updatepcg_alpha"][0])
    rdotr = gpu.block_update_pcg(w, Ap, alpha, r)
    # the elementwise update is one multiply-add per entry: exact up to contraction
    assert np.max(np.abs(r.cpu().numpy() - g["updatepcg_r_out"])) <= 2 * EPS * np.max(np.abs(g["updatepcg_r_out"]))
    ref = float(g["updatepcg_rdotr"][0])
    assert abs(rdotr - ref) <= TOL * ref


@pytest.mark.parametrize("N", [1, 255, 4096, 100003])
def test_block_update_pcg_three_fields(gpu, torch_mod, N):
    torch = torch_mod
    off = field_offset(N)
    rng = np.random.default_rng(N)
    r = rng.uniform(-1, 1, 3 * off)
    Ap = rng.uniform(-1, 1, 3 * off)
    w = rng.uniform(0, 1, N)
    alpha = 0.37
    rn = r.copy()
    for f in range(3):
        s = slice(f * off, f * off + N)
        rn[s] = r[s] - alpha * Ap[s]
    ref = math.fsum((w[None] * rn.reshape(3, off)[:, :N] ** 2).ravel().tolist())
    rd = torch.from_numpy(r).cuda()
    got = gpu.block_update_pcg(torch.from_numpy(w).cuda(), torch.from_numpy(Ap).cuda(), alpha, rd,
                               N=N, nfields=3, offset=off)
    assert abs(got - ref) <= TOL * ref
    out = rd.cpu().numpy()
    assert np.max(np.abs(out - rn)) <= 2 * EPS
    # padding between the fields is untouched
    for f in range(3):
        assert np.array_equal(out[f * off + N:(f + 1) * off], r[f * off + N:(f + 1) * off])
