"""Solves on a mesh read from a .re2 file whose elements have their corners in
randomly rotated orders (written by oracle/re2.py), against the same box from
the structured mesher: the assembled problem is the same, so Jacobi-PCG takes
the same iterations (+-1) and the solutions, both solved to 1e-12, agree node
by node (matched by coordinates) to 1e-8; multigrid with Schwarz (RAS) smoothing, whose face
exchange must follow each element's orientation, converges in the same number
of iterations (+-2) to the same solution within 2e-7: its FP32 prolongation is
elementwise, as in the reference (ellipticMultiGridLevel.cpp:70-83, no
gather-scatter after it), so elements of different orientation round the
values of a shared node differently, and the FP64 solution inherits those
FP32-rounding-level jumps (measured 3.3e-8)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rotated_re2(tmp_path, nx, ny, nz, seed):
    import oracle.re2 as r2
    c = r2.box_corners(nx, ny, nz)
    rots = r2.rotations()
    rng = np.random.default_rng(seed)
    for e in range(c.shape[0]):
        c[e] = c[e][rots[rng.integers(0, 24)]]
    codes = {}
    for e in range(c.shape[0]):
        for f, cs in enumerate(r2.FILE_FACE_CORNERS):
            p = c[e][list(cs)]
            for ax in range(3):
                if np.all(np.abs(p[:, ax]) < 1e-12) or np.all(np.abs(p[:, ax] - 1) < 1e-12):
                    codes[(e, f)] = "O  "
    path = tmp_path / "rot.re2"
    r2.write(path, c, codes)
    return path


def solve(gpu, torch, mesh, f, **kw):
    el = gpu.Elliptic(mesh, tol=1e-12, max_iter=2000, **kw)
    gw = el.ggeo.cpu().numpy().reshape(-1, 7, mesh.Np)[:, 6].ravel()
    x = torch.zeros(mesh.Nlocal, dtype=torch.float64, device="cuda")
    st = el.solve(torch.from_numpy(gw * f(mesh.x, mesh.y, mesh.z)).cuda(), x)
    return x.cpu().numpy(), st


def by_coordinates(mesh, v):
    key = np.round(np.stack([mesh.x, mesh.y, mesh.z], 1) * 1e9).astype(np.int64)
    _, first = np.unique(key, axis=0, return_index=True)
    order = np.lexsort(key[first].T[::-1])
    return v[first][order]


@pytest.mark.parametrize("kw", [dict(precon="jacobi", solver="pcg"),
                                dict(precon="multigrid", solver="fcg", levels=[5, 3, 1],
                                     smoother="chebyshev+ras", cheby_degree=2, coarse="smooth_jacobi")])
def test_rotated_re2_mesh_solves_like_the_box(gpu, torch_mod, tmp_path, kw):
    torch = torch_mod
    nx, ny, nz, N = 4, 3, 3, 5
    path = rotated_re2(tmp_path, nx, ny, nz, 7)
    m = gpu.Re2Mesh(path, N)
    b = gpu.BoxMesh(nx, ny, nz, N)
    assert np.array_equal(np.sort(m.EToB), np.sort(b.EToB))
    f = lambda x, y, z: np.sin(np.pi * x) * np.sin(2 * np.pi * y) * (1 + z)
    xm, sm = solve(gpu, torch, m, f, **kw)
    xb, sb = solve(gpu, torch, b, f, **kw)
    print(kw["precon"], "iterations re2", sm["iterations"], "box", sb["iterations"], flush=True)
    assert sm["converged"] and sb["converged"]
    tol_it = 1 if kw["precon"] == "jacobi" else 2
    assert abs(sm["iterations"] - sb["iterations"]) <= tol_it
    um, ub = by_coordinates(m, xm), by_coordinates(b, xb)
    assert um.shape == ub.shape
    bar = 1e-8 if kw["precon"] == "jacobi" else 2e-7
    assert np.max(np.abs(um - ub)) <= bar * np.max(np.abs(ub))
