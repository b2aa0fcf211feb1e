"""Periodic connectivity of .re2 meshes ('P' boundary records: bc(1) = the
connected element, bc(2) = its face, reader_re2.f:503-541): files written by
oracle/re2.py for a box periodic in x (and in x and z), elements with their
corners in random rotated orders.  Host checks: the nodes of paired faces share
global ids exactly when they share coordinates modulo the period, and the
number of ids is that of the periodic box; one or two elements across a
periodic direction are refused.  GPU: Jacobi-PCG on the periodic mesh against
the oracle's PCG on the same ids and geometry (iterations +-1, solution 1e-8),
and the manufactured periodic solution sin(2 pi x) sin(pi y) sin(pi z)."""
import numpy as np
import pytest

import oracle.re2 as r2


def periodic_re2(path, nx, ny, nz, seed, periodic=("x",), rotate=True):
    """A box periodic in the given directions, 'O  ' on the other boundary faces."""
    c = r2.box_corners(nx, ny, nz)
    rng = np.random.default_rng(seed)
    if rotate:
        rots = r2.rotations()
        for e in range(c.shape[0]):
            c[e] = c[e][rots[rng.integers(0, 24)]]
    E = c.shape[0]
    n = (nx, ny, nz)

    def face_on(e, ax, val):
        for f, cs in enumerate(r2.FILE_FACE_CORNERS):
            if np.all(np.abs(c[e][list(cs)][:, ax] - val) < 1e-12):
                return f
        return None

    def cell(e):
        ctr = c[e].mean(0)
        return tuple(int(np.floor(ctr[a] * n[a])) for a in range(3))

    where = {cell(e): e for e in range(E)}
    codes, params = {}, {}
    for e in range(E):
        ijk = cell(e)
        for ax, name in enumerate("xyz"):
            for val in (0.0, 1.0):
                f = face_on(e, ax, val)
                if f is None:
                    continue
                if name in periodic:
                    other = list(ijk)
                    other[ax] = n[ax] - 1 if val == 0.0 else 0
                    e2 = where[tuple(other)]
                    f2 = face_on(e2, ax, 1.0 - val)
                    codes[(e, f)] = "P  "
                    params[(e, f)] = (e2 + 1, f2 + 1)
                else:
                    codes[(e, f)] = "O  "
    r2.write(path, c, codes, params=params)
    return path


def periodic_key(m, periodic):
    x = np.stack([m.x, m.y, m.z], 1).copy()
    for ax, name in enumerate("xyz"):
        if name in periodic:
            x[:, ax] = np.mod(np.round(x[:, ax], 9), 1.0)
    return np.round(x * 1e8).astype(np.int64)


@pytest.mark.parametrize("periodic,shape", [(("x",), (4, 3, 3)), (("x", "z"), (3, 3, 4))])
def test_periodic_ids_follow_coordinates(tmp_path, periodic, shape):
    import nekrs_amd as gpu
    N = 4
    path = periodic_re2(tmp_path / "p.re2", *shape, seed=3, periodic=periodic)
    m = gpu.Re2Mesh(path, N)
    nx, ny, nz = shape
    per = {"x": nx * N, "y": ny * N, "z": nz * N}
    count = 1
    for a in "xyz":
        count *= per[a] if a in periodic else per[a] + 1
    assert len(np.unique(m.gid)) == count
    _, ik = np.unique(periodic_key(m, periodic), axis=0, return_inverse=True)
    _, ig = np.unique(m.gid, return_inverse=True)
    pairs = set(zip(ik.ravel().tolist(), ig.ravel().tolist()))
    assert len(pairs) == ik.max() + 1 == ig.max() + 1          # a bijection
    assert np.count_nonzero(m.EToB) == sum(2 * (shape[(a + 1) % 3] * shape[(a + 2) % 3])
                                           for a, nm in enumerate("xyz") if nm not in periodic)


@pytest.mark.parametrize("nx", [1, 2])
def test_too_few_elements_across_a_period_are_refused(tmp_path, nx):
    import nekrs_amd as gpu
    path = periodic_re2(tmp_path / "p.re2", nx, 2, 2, seed=1, rotate=False)
    with pytest.raises(gpu.NrsError):
        gpu.Re2Mesh(path, 3)


def oracle_problem(orc, m, shape):
    """The oracle's elliptic on the same nodes, ids and Dirichlet faces."""
    mo = orc.BoxMesh(*shape, m.N)
    mo.x, mo.y, mo.z = m.x.copy(), m.y.copy(), m.z.copy()
    mo.gid = m.gid.copy()
    mo.EToB = m.EToB.astype(np.int32).copy()
    _, wq = orc.gll(m.N)
    mo.ggeo = np.zeros(mo.E * 7 * mo.Np)
    orc.lib().orc_geofacs(m.N, mo.E, mo.D.ravel().copy(), wq, mo.x, mo.y, mo.z, mo.ggeo, None)
    return orc.Elliptic(mo)


@pytest.mark.gpu
def test_periodic_solve_vs_oracle_and_exact(gpu, torch_mod, orc, tmp_path):
    import oracle.mg as omg
    torch = torch_mod
    shape, N = (4, 3, 3), 5
    path = periodic_re2(tmp_path / "p.re2", *shape, seed=5)
    m = gpu.Re2Mesh(path, N)
    el = gpu.Elliptic(m, solver="pcg", precon="jacobi", tol=1e-10, max_iter=2000)
    gw = el.ggeo.cpu().numpy().reshape(-1, 7, m.Np)[:, 6].ravel()
    u = np.sin(2 * np.pi * m.x) * np.sin(np.pi * m.y) * np.sin(np.pi * m.z)
    rhs = gw * 6 * np.pi ** 2 * u
    x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    st = el.solve(torch.from_numpy(rhs).cuda(), x)
    eo = oracle_problem(orc, m, shape)
    inv = eo.inv_diag()
    xo, info = omg.fcg_solve(eo, rhs, lambda v: inv * v, flexible=False, tol=1e-10, maxit=2000)
    xg = x.cpu().numpy()
    err = np.max(np.abs(xg - u))
    print("periodic: iterations", st["iterations"], "oracle", info["iterations"], "max error vs exact",
          err, flush=True)
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert np.max(np.abs(xg - xo)) <= 1e-8 * np.max(np.abs(xo))
    assert err <= 1e-3
