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


def test_curved_re2_mesh_vs_oracle(gpu, torch_mod, tmp_path):
    """A .re2 box whose elements (in rotated corner orders) have midside-node
    curved edges on a quadratic deformation of the box: the operator on the
    library's curved coordinates agrees with the oracle's on its own
    restatement of the curved map (1e-12), and Jacobi-PCG takes the oracle's
    iterations (+-1) to the same solution (1e-8)."""
    import oracle
    import oracle.re2 as r2
    torch = torch_mod
    nx, ny, nz, N = 3, 3, 2, 6
    c = r2.box_corners(nx, ny, nz)
    rots = r2.rotations()
    rng = np.random.default_rng(11)
    for e in range(c.shape[0]):
        c[e] = c[e][rots[rng.integers(0, 24)]]

    def phi(p):
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        return np.stack([x + 0.2 * x * (1 - x) * (1 + y), y - 0.15 * y * (1 - y) * (1 + z),
                         z + 0.2 * z * (1 - z) * (1 + x)], -1)

    recs = []
    for e in range(c.shape[0]):
        for k, (i, j, kk) in enumerate(r2.EDGE_MID):
            ends = [(t if i == 1 else i) // 2 + 2 * ((t if j == 1 else j) // 2) + 4 * ((t if kk == 1 else kk) // 2)
                    for t in (0, 2)]
            mid = phi(0.5 * (c[e, ends[0]] + c[e, ends[1]]))
            recs.append((e, k + 1, (mid[0], mid[1], mid[2], 0.0, 0.0), "m"))
    cp = phi(c)
    codes = {}
    for e in range(c.shape[0]):
        for f, cs in enumerate(r2.FILE_FACE_CORNERS):
            p = c[e][list(cs)]
            for ax in range(3):
                if np.all(np.abs(p[:, ax]) < 1e-12) or np.all(np.abs(p[:, ax] - 1) < 1e-12):
                    codes[(e, f)] = "O  "
    path = tmp_path / "curved.re2"
    r2.write(path, cp, codes, curves=recs)
    m = gpu.Re2Mesh(path, N)
    print("curved re2: %d elements, %d curve records" % (m.E, m.ncurve), flush=True)
    xo, yo, zo = m.x.copy(), m.y.copy(), m.z.copy()
    # the oracle's own curved coordinates: trilinear from the corners, then its midside map
    vx, vy, vz = (np.ascontiguousarray(cp[:, :, d].ravel()) for d in range(3))
    xs, ys, zs, _ = gpu.mesh_from_vertices(N, vx, vy, vz)
    r2.midside_map(N, cp, recs, xs, ys, zs)
    assert max(np.max(np.abs(xs - xo)), np.max(np.abs(ys - yo)), np.max(np.abs(zs - zo))) < 1e-14
    om = oracle.ArrayMesh(N, xs, ys, zs, m.gid, m.EToB)
    eo = oracle.Elliptic(om)
    el = gpu.Elliptic(m, precon="jacobi", tol=1e-8, max_iter=2000)
    q = np.random.default_rng(3).uniform(-1, 1, m.Nlocal)
    Aq = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    el.operator(torch.from_numpy(q).cuda(), Aq)
    ref = eo.operator(q)
    err = np.max(np.abs(Aq.cpu().numpy() - ref)) / np.max(np.abs(ref))
    print("operator rel err %.2e" % err, flush=True)
    assert err < 1e-12
    gw = om.ggeo.reshape(-1, 7, om.Np)[:, 6].ravel()
    rhs = gw * np.sin(np.pi * xs) * np.cos(np.pi * ys) * (1 + zs)
    xo_sol, info = eo.solve(rhs, tol=1e-8)
    x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    st = el.solve(torch.from_numpy(rhs).cuda(), x)
    print("iterations gpu %d oracle %d" % (st["iterations"], info["iterations"]), flush=True)
    assert st["converged"] and abs(st["iterations"] - info["iterations"]) <= 1
    assert np.max(np.abs(x.cpu().numpy() - xo_sol)) <= 1e-8 * np.max(np.abs(xo_sol))
