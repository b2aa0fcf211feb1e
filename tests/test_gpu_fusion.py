"""The operator face fusion (N = 7: the shared face of consecutive elements
meeting i = N to i = 0 summed inside the Ax kernel, those nodes left out of the
operator's gather-scatter) gives the unfused operator bit for bit: masked and
unmasked, every chunk variant, on the structured box, on a box whose element
row length leaves partial blocks, and on a .re2 mesh with rotated elements
(where only identically oriented faces fuse); solves match iteration for
iteration and bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def pair(gpu, mesh, **kw):
    """Fused and unfused handles; autotune off, so the fusion is kept whatever it times."""
    on = gpu.Elliptic(mesh, operatorFusion=1, autotune=0, **kw)
    off = gpu.Elliptic(mesh, operatorFusion=0, autotune=0, **kw)
    return on, off


@pytest.mark.parametrize("shape", [(8, 3, 2), (6, 2, 3), (5, 5, 2)])
def test_operator_fused_equals_unfused(gpu, torch_mod, shape):
    torch = torch_mod
    m = gpu.BoxMesh(*shape, 7, "kershaw", 0.3)
    on, off = pair(gpu, m, precon="jacobi")
    assert off.fused_pairs == 0
    assert on.fused_pairs > 0
    print(shape, "fused pairs", on.fused_pairs, flush=True)
    q = torch.from_numpy(np.random.default_rng(1).uniform(-1, 1, m.Nlocal)).cuda()
    for masked in (True, False):
        a, b = torch.empty_like(q), torch.empty_like(q)
        off.operator(q, b, masked=masked)
        for chunks in (1, 2, 4, 8, 16):
            gpu.set_variant("op", 7, "f64", m.E, chunks)
            on.operator(q, a, masked=masked)
            assert torch.equal(a, b), (masked, chunks)
        gpu.set_variant("op", 7, "f64", m.E, 0)


def test_helmholtz_and_solve_bitwise(gpu, torch_mod):
    torch = torch_mod
    m = gpu.BoxMesh(8, 4, 3, 7, "kershaw", 0.25)
    kw = dict(lambda0=1.0, lambda1=3.0, poisson=False, precon="multigrid", solver="fcg",
              levels=[7, 3, 1], smoother="chebyshev+jacobi", mgPrecision=1, coarse="smooth_jacobi",
              tol=1e-10, max_iter=200)
    on, off = pair(gpu, m, **kw)
    rhs = np.random.default_rng(2).uniform(-1, 1, m.Nlocal)
    xs, sts = [], []
    for el in (on, off):
        x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
        sts.append(el.solve(torch.from_numpy(rhs).cuda(), x))
        xs.append(x)
    assert sts[0]["iterations"] == sts[1]["iterations"] and sts[0]["converged"]
    assert torch.equal(xs[0], xs[1])


def test_rotated_mesh_partial_fusion(gpu, torch_mod, tmp_path):
    import oracle.re2 as r2
    torch = torch_mod
    c = r2.box_corners(8, 2, 2)
    rots = r2.rotations()
    rng = np.random.default_rng(3)
    for e in range(0, c.shape[0], 3):                # every third element rotated
        c[e] = c[e][rots[rng.integers(0, 24)]]
    codes = {}
    for e in range(c.shape[0]):
        for f, cs in enumerate(r2.FILE_FACE_CORNERS):
            p = c[e][list(cs)]
            if any(np.all(np.abs(p[:, ax] - v) < 1e-12) for ax in range(3) for v in (0.0, 1.0)):
                codes[(e, f)] = "O  "
    path = tmp_path / "part.re2"
    r2.write(path, c, codes)
    m = gpu.Re2Mesh(path, 7)
    on, off = pair(gpu, m, precon="jacobi")
    full = gpu.Elliptic(gpu.BoxMesh(8, 2, 2, 7), precon="jacobi", operatorFusion=1, autotune=0)
    assert 0 < on.fused_pairs < full.fused_pairs
    q = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, m.Nlocal)).cuda()
    a, b = torch.empty_like(q), torch.empty_like(q)
    on.operator(q, a)
    off.operator(q, b)
    assert torch.equal(a, b)


def test_autotune_keeps_fusion_only_when_faster(gpu, torch_mod):
    """With autotune on, the setup times both forms; whichever it keeps, the operator
    equals the unfused one bit for bit."""
    torch = torch_mod
    m = gpu.BoxMesh(8, 4, 4, 7, "kershaw", 0.3)
    tuned = gpu.Elliptic(m, precon="jacobi", operatorFusion=1, autotune=1)
    off = gpu.Elliptic(m, precon="jacobi", operatorFusion=0, autotune=0)
    print("fused pairs kept:", tuned.fused_pairs, flush=True)
    q = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, m.Nlocal)).cuda()
    a, b = torch.empty_like(q), torch.empty_like(q)
    tuned.operator(q, a)
    off.operator(q, b)
    assert torch.equal(a, b)


@pytest.mark.parametrize("shape", [(8, 3, 2), (5, 5, 2)])
def test_operator_fused_blocks_of_two_equals_unfused(gpu, torch_mod, shape):
    """operatorFusion = 2: the fused Ax kernel with two elements per block (pairs
    e, e+1 for even e) gives the unfused operator bit for bit at every chunk count."""
    torch = torch_mod
    m = gpu.BoxMesh(*shape, 7, "kershaw", 0.3)
    two = gpu.Elliptic(m, precon="jacobi", autotune=0, operatorFusion=2)
    four = gpu.Elliptic(m, precon="jacobi", autotune=0, operatorFusion=1)
    off = gpu.Elliptic(m, precon="jacobi", autotune=0, operatorFusion=0)
    print(shape, "fused pairs: blocks of two", two.fused_pairs, "of four", four.fused_pairs, flush=True)
    assert 0 < two.fused_pairs < four.fused_pairs
    q = torch.from_numpy(np.random.default_rng(2).uniform(-1, 1, m.Nlocal)).cuda()
    for masked in (True, False):
        a, b = torch.empty_like(q), torch.empty_like(q)
        off.operator(q, b, masked=masked)
        for chunks in (1, 4, 16):
            gpu.set_variant("op", 7, "f64", m.E, chunks)
            two.operator(q, a, masked=masked)
            assert torch.equal(a, b), (masked, chunks)
        gpu.set_variant("op", 7, "f64", m.E, 0)


def test_third_copy_keeps_pair_unfused(gpu, torch_mod):
    """A face-interior id of a fusable pair that also appears a third time
    elsewhere (caller-supplied global ids may do that) keeps that pair out of
    the fusion: the fused operator still equals the unfused one bit for bit,
    and one fewer pair is fused than on the plain box."""
    torch = torch_mod
    m = gpu.BoxMesh(8, 2, 2, 7)
    plain = gpu.Elliptic(m, precon="jacobi", operatorFusion=1, autotune=0)
    # element 1's face i = N against element 2's face i = 0 (a fused pair of the
    # first block of four): give node (i=7, j=3, k=3) of element 1 a third copy in
    # the interior of element 6 (an interior node is otherwise a singleton)
    a = 1 * 512 + 3 * 64 + 3 * 8 + 7
    other = 6 * 512 + 4 * 64 + 4 * 8 + 4
    m.gid[other] = m.gid[a]
    on, off = pair(gpu, m, precon="jacobi")
    print("fused pairs: plain %d, third copy %d" % (plain.fused_pairs, on.fused_pairs), flush=True)
    assert on.fused_pairs == plain.fused_pairs - 1
    q = torch.from_numpy(np.random.default_rng(5).uniform(-1, 1, m.Nlocal)).cuda()
    x, y = torch.empty_like(q), torch.empty_like(q)
    on.operator(q, x)
    off.operator(q, y)
    assert torch.equal(x, y)
    # the third copy took part in the sum
    assert x[a].item() == x[other].item()
