"""Multigrid level operators that form their geometric factors from the element
corners (mgGeometry = 1, the default, on meshes of trilinear elements) against
the same operators streaming the stored factors (mgGeometry = 0, the
reference's way): every level's A x agrees to 1e-13 relative in FP64 and 1e-5 in
FP32 (there the stored factors are FP64 values rounded once, the formed ones
FP32 arithmetic throughout), at orders 1..9, Poisson and Helmholtz, on the
Kershaw box and a .re2 box with rotated elements; on a curved (sine-mapped)
box the option changes nothing (bit for bit); and full p-MG solves take the
same iterations (+-1) to the same solution (1e-8)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def handles(gpu, mesh, levels, prec, **kw):
    common = dict(precon="multigrid", solver="fcg", levels=levels, smoother="chebyshev+jacobi",
                  cheby_degree=2, coarse="smooth_jacobi", mgPrecision=prec, tol=1e-8, max_iter=300, **kw)
    return (gpu.Elliptic(mesh, mgGeometry=1, **common), gpu.Elliptic(mesh, mgGeometry=0, **common))


def level_ax_rel(torch, on, off, nlev, dtype, seed):
    rel = []
    for lv in range(nlev):
        n = on.mg_info()[0][lv]["Nlocal"]
        q = torch.from_numpy(np.random.default_rng(seed + lv).uniform(-1, 1, n)).to(dtype).cuda()
        a, b = torch.empty_like(q), torch.empty_like(q)
        on.mg_level(lv, "ax", q, a)
        off.mg_level(lv, "ax", q, b)
        torch.cuda.synchronize()
        rel.append(float(torch.max(torch.abs(a.double() - b.double())) / torch.max(torch.abs(b.double()))))
    return rel


@pytest.mark.parametrize("N", [2, 3, 5, 7, 9])
@pytest.mark.parametrize("prec", [0, 1])
def test_corner_geometry_level_operators(gpu, torch_mod, N, prec):
    torch = torch_mod
    levels = [N] + ([N // 2 + 1] if N > 3 else []) + [1]
    levels = sorted(set(levels), reverse=True)
    mesh = gpu.BoxMesh(4, 3, 3, N, "kershaw", 0.3)
    for kw in ({}, {"lambda0": 1.0, "lambda1": 3.0, "poisson": False}):
        on, off = handles(gpu, mesh, levels, prec, **kw)
        dtype = torch.float64 if prec == 1 else torch.float32
        rel = level_ax_rel(torch, on, off, len(levels), dtype, 10 * N)
        print("N=%d prec=%d %s levels %s rel %s" % (N, prec, "helmholtz" if kw else "poisson", levels, rel),
              flush=True)
        bar = 1e-13 if prec == 1 else 1e-5
        assert max(rel) < bar, rel


def test_curved_mesh_unchanged(gpu, torch_mod):
    """Sine-mapped nodes are not trilinear: the option has no effect."""
    torch = torch_mod
    mesh = gpu.BoxMesh(3, 3, 2, 5, "sine", 0.0)
    on, off = handles(gpu, mesh, [5, 3, 1], 0)
    rel = level_ax_rel(torch, on, off, 3, torch.float32, 1)
    assert max(rel) == 0.0, rel


@pytest.mark.parametrize("prec", [0, 1])
def test_solves_match(gpu, torch_mod, prec):
    torch = torch_mod
    mesh = gpu.BoxMesh(6, 5, 5, 7, "kershaw", 0.3)
    on, off = handles(gpu, mesh, [7, 3, 1], prec)
    rhs = torch.from_numpy(np.random.default_rng(4).uniform(-1, 1, mesh.Nlocal)).cuda()
    xa, xb = torch.zeros_like(rhs), torch.zeros_like(rhs)
    sa = on.solve(rhs.clone(), xa)
    sb = off.solve(rhs.clone(), xb)
    print("prec %d iterations corner %d stored %d" % (prec, sa["iterations"], sb["iterations"]), flush=True)
    assert sa["converged"] and sb["converged"]
    assert abs(sa["iterations"] - sb["iterations"]) <= 1
    assert float(torch.max(torch.abs(xa - xb)) / torch.max(torch.abs(xb))) < 1e-8


def test_rotated_re2_mesh(gpu, torch_mod, tmp_path):
    """Elements whose corners come in rotated orders (the map's axes differ per element)."""
    import oracle.re2 as r2
    torch = torch_mod
    c = r2.box_corners(3, 3, 2)
    rots = r2.rotations()
    rng = np.random.default_rng(5)
    for e in range(c.shape[0]):
        c[e] = c[e][rots[rng.integers(0, 24)]]
    c[..., 0] += 0.1 * c[..., 1] * c[..., 2]   # a trilinear (not affine) distortion
    path = tmp_path / "rot.re2"
    r2.write(path, c)
    mesh = gpu.Re2Mesh(path, 7)
    on, off = handles(gpu, mesh, [7, 3, 1], 1)
    rel = level_ax_rel(torch, on, off, 3, torch.float64, 2)
    assert max(rel) < 1e-13, rel
