"""Pinning of the pressure right-hand side oracle (oracle/tombo.py) by what
the mathematics fixes, on affine meshes (a sheared, rotated, scaled box) where
polynomial fields of degree <= N are represented exactly: the metric terms
invert the map's Jacobian; normals are unit and outward and the WSJ of a face
sum to its area; curl and gradient of quadratics are exact; the weak
divergence integrates by parts against a linear test function; the surface
terms of U = (x, 0, 0) on inflow faces sum to -coef * volume and those of a
constant field to zero; the whole chain reproduces closed forms for
curl curl of a quadratic field."""
import numpy as np
import pytest

A = np.array([[1.3, 0.2, -0.1], [0.1, 0.9, 0.25], [-0.2, 0.15, 1.1]])
B = np.array([0.3, -0.2, 0.5])


def affine_mesh(orc, nx=3, ny=2, nz=2, N=4):
    m = orc.BoxMesh(nx, ny, nz, N)
    X = np.stack([m.x, m.y, m.z])
    Y = A @ X + B[:, None]
    return m, Y[0].copy(), Y[1].copy(), Y[2].copy()


def geom(orc, N=4):
    import oracle.tombo as ot
    m, x, y, z = affine_mesh(orc, N=N)
    return m, x, y, z, ot.Geometry(N, x, y, z, m.gid)


def test_metrics_invert_the_map(orc):
    m, x, y, z, g = geom(orc)
    h = np.array([1 / 3, 1 / 2, 1 / 2])                  # element sizes of the unit box
    dXdr = A * (h / 2)[None, :]                          # columns: d(x,y,z)/d(r,s,t)
    M = np.stack([np.stack([g.rx, g.ry, g.rz], -1), np.stack([g.sx, g.sy, g.sz], -1),
                  np.stack([g.tx, g.ty, g.tz], -1)], -2)  # [...][3 (r,s,t)][3 (x,y,z)]
    prod = M @ dXdr
    assert np.max(np.abs(prod - np.eye(3))) <= 1e-12
    vol = abs(np.linalg.det(A))
    assert abs(g.JW.sum() - vol) <= 1e-12 * vol


def test_normals_outward_unit_and_areas(orc):
    m, x, y, z, g = geom(orc)
    Nq = g.Nq
    X = np.stack([x, y, z]).reshape(3, g.E, Nq, Nq, Nq)
    centre = X.mean(axis=(2, 3, 4))
    for f in range(6):
        sel, nx, ny, nz, WSJ = g.face(f)
        assert np.max(np.abs(nx ** 2 + ny ** 2 + nz ** 2 - 1)) <= 1e-13
        fc = np.stack([X[c][sel].mean(axis=(1, 2)) for c in range(3)])
        d = (fc - centre)
        dot = nx.mean(axis=(1, 2)) * d[0] + ny.mean(axis=(1, 2)) * d[1] + nz.mean(axis=(1, 2)) * d[2]
        assert np.all(dot > 0)
        # area of face f of every element from the cross product of its edge vectors
        e1 = {0: 0, 1: 0, 2: 1, 3: 0, 4: 1, 5: 0}[f]
        e2 = {0: 1, 1: 2, 2: 2, 3: 2, 4: 2, 5: 1}[f]
        h = np.array([1 / 3, 1 / 2, 1 / 2])
        area = np.linalg.norm(np.cross(A[:, e1] * h[e1], A[:, e2] * h[e2]))
        assert np.max(np.abs(WSJ.sum(axis=(1, 2)) - area)) <= 1e-12


def test_curl_and_gradient_exact(orc):
    m, x, y, z, g = geom(orc)
    L = x.size
    U = np.concatenate([y * y, z * z, x * x])
    W = g.curl(U) / np.tile(g.JW.ravel(), 3)
    ref = np.concatenate([-2 * z, -2 * x, -2 * y])
    assert np.max(np.abs(W - ref)) <= 1e-10
    p = x * y + z * z
    G = g.gradient(p) / np.tile(g.JW.ravel(), 3)
    ref = np.concatenate([y, x, 2 * z])
    assert np.max(np.abs(G - ref)) <= 1e-10


def test_weak_divergence_by_parts(orc):
    """sum_i phi_i wdiv(u)_i = sum JW grad(phi).u for phi = a linear function."""
    m, x, y, z, g = geom(orc)
    rng = np.random.default_rng(1)
    U = rng.standard_normal(3 * x.size)
    wd = g.wdiv(U)
    JW = np.tile(g.JW.ravel(), 1)
    for phi, grad in ((x, (1, 0, 0)), (y, (0, 1, 0)), (2 * x - z, (2, 0, -1))):
        lhs = np.sum(phi * wd)
        L = x.size
        rhs = np.sum(JW * (grad[0] * U[:L] + grad[1] * U[L:2 * L] + grad[2] * U[2 * L:]))
        assert abs(lhs - rhs) <= 1e-11 * np.sum(np.abs(JW * U[:L]))


def test_surface_flux_sums(orc):
    m, x, y, z, g = geom(orc)
    L = x.size
    bc = np.where(m.EToB != 0, 3, 0)                     # every boundary face an inflow
    zero = np.zeros(3 * L)
    vol = abs(np.linalg.det(A))
    s = g.surface(bc, 2.0, zero, np.concatenate([x, 0 * x, 0 * x]), L)
    assert abs(s.sum() + 2.0 * vol) <= 1e-12
    s = g.surface(bc, 2.0, zero, np.concatenate([1 + 0 * x, 2 + 0 * x, -1 + 0 * x]), L)
    assert abs(s.sum()) <= 1e-12
    # symmetry faces take n.F instead, walls nothing
    bcs = np.where(m.EToB != 0, 7, 0)
    s = g.surface(bcs, 2.0, np.concatenate([x, 0 * x, 0 * x]), zero, L)
    assert abs(s.sum() + vol) <= 1e-12
    assert np.all(g.surface(np.where(m.EToB != 0, 1, 0), 2.0, U := np.ones(3 * L), U, L) == 0)


@pytest.mark.parametrize("visc", [True, False])
def test_pressure_rhs_closed_form(orc, visc):
    """Walls everywhere (no surface terms), div = c + a.x: with Ue =
    (y^2, z^2, x^2), curl curl Ue = (-2, -2, -2), BF = JW f with f = (x, 1, z), and
    phi = x: sum_i x_i rhs_i = sum JW (f_x + visc nu (2 + 4/3 a_x)) + g0idt sum x JW div."""
    m, x, y, z, g = geom(orc)
    L = x.size
    bc = np.where(m.EToB != 0, 1, 0)
    Ue = np.concatenate([y * y, z * z, x * x])
    JW = g.JW.ravel()
    BF = np.concatenate([JW * x, JW * 1.0, JW * z])
    nu_m, nu_r = 0.7, 1.5
    mue = np.full(L, nu_m)
    irho = np.full(L, nu_r)
    c, ax, ay = 0.3, 0.5, -0.2
    div = c + ax * x + ay * y
    g0idt = 4.0
    rhs = g.pressure_rhs(bc, Ue, np.zeros(3 * L), BF, mue, irho, div, g0idt, visc)
    nu = nu_m * nu_r
    expect = (np.sum(JW * (x + (nu * (2 + 4.0 / 3.0 * ax) if visc else 0.0))) +
              g0idt * np.sum(x * JW * div))
    assert abs(np.sum(x * rhs) - expect) <= 1e-10 * abs(expect)
    expect_y = (np.sum(JW * (1.0 + (nu * (2 + 4.0 / 3.0 * ay) if visc else 0.0))) +
                g0idt * np.sum(y * JW * div))
    assert abs(np.sum(y * rhs) - expect_y) <= 1e-10 * abs(expect_y)
    # phi = 1: the weak divergence sums to zero
    assert abs(np.sum(rhs) - g0idt * np.sum(JW * div)) <= 1e-10 * abs(g0idt * np.sum(JW * div))


def test_stress_term_closed_form(orc):
    """The stress-formulation term (pressureStressHex3D.okl) of quadratic
    velocity and viscosity fields on the affine mesh, where the nodal
    derivatives are exact: JW (2 sym(grad U) grad(mu)/mu - 2/3 div grad(mu)/mu)
    against the closed form; a constant viscosity gives zero."""
    m, x, y, z, g = geom(orc)
    L = x.size
    U = np.concatenate([x * y + z, y * z - 0.5 * x * x, z * x + y * y])
    mu = 1.5 + 0.2 * x - 0.1 * y * y + 0.05 * x * z
    div = 0.3 + 0.7 * x - 0.2 * z                          # any field: the term is linear in it
    S = g.stress(U, mu, div).reshape(3, L)
    grad_u = np.stack([np.stack([y, x, np.ones(L)]),            # d(u)/d(x, y, z)
                       np.stack([-x, z, y]),                     # d(v)/d(...)
                       np.stack([z, 2 * y, x])])                 # d(w)/d(...)
    gm = np.stack([0.2 + 0.05 * z, -0.2 * y, 0.05 * x]) / mu
    sym = grad_u + grad_u.transpose(1, 0, 2)
    w = np.einsum("abn,bn->an", sym, gm)
    ref = g.JW.ravel()[None, :] * (w - (2.0 / 3.0) * div[None, :] * gm)
    assert np.max(np.abs(S - ref)) <= 1e-10 * np.max(np.abs(ref))
    assert np.max(np.abs(g.stress(U, np.full(L, 2.0), div))) <= 1e-12
    # in the right-hand side it is taken from the second curl
    rng = np.random.default_rng(0)
    bc = np.zeros(6 * g.E, dtype=np.int32)
    BF = rng.uniform(-1, 1, 3 * L)
    irho = np.full(L, 0.7)
    a = g.pressure_rhs(bc, U, U, BF, mu, irho, div, 2.0, visc=True, stress=True)
    b = g.pressure_rhs(bc, U, U, BF, mu, irho, div, 2.0, visc=True, stress=False)
    nu = np.tile(mu * irho, 3)
    dU = g.ogs.gs(nu * S.ravel(), "add", nvec=3, stride=L) * np.tile(g.invLMM, 3)
    assert np.max(np.abs((a - b) - g.wdiv(dU))) <= 1e-10 * np.max(np.abs(a))
