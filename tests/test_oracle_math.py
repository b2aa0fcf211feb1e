"""The oracle against what the mathematics fixes: closed forms, invariants,
brute force on tiny inputs, and textbook CG on an explicitly assembled matrix.
Each check is chosen so that a plausible slip (transposed D, swapped G index,
dropped term, wrong row order, wrong sign) fails it."""
import numpy as np
import numpy.polynomial.legendre as npl
import pytest


@pytest.mark.parametrize("N", range(1, 12))
def test_gll_closed_form(orc, N):
    z, w = orc.gll(N)
    interior = np.sort(npl.Legendre.basis(N).deriv().roots().real)
    assert np.allclose(z[1:-1], interior, atol=1e-14, rtol=0)
    assert z[0] == -1.0 and z[-1] == 1.0
    PN = npl.legval(z, [0] * N + [1])
    assert np.allclose(w, 2.0 / (N * (N + 1) * PN ** 2), rtol=1e-14)
    for k in range(2 * N):        # exact for degree <= 2N-1
        exact = (1 - (-1) ** (k + 1)) / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-13


@pytest.mark.parametrize("N", range(1, 12))
def test_dmatrix_exact_on_polynomials(orc, N):
    z, _ = orc.gll(N)
    D = orc.dmatrix(N)
    for k in range(N + 1):
        dk = k * z ** (k - 1) if k > 0 else np.zeros_like(z)
        assert np.allclose(D @ z ** k, dk, atol=1e-11 * max(1, k * k))


def test_interp_exact(orc):
    for Nin, Nout in [(3, 7), (7, 3), (1, 7), (5, 9)]:
        zin, _ = orc.gll(Nin)
        zout, _ = orc.gll(Nout)
        I = orc.interp(zin, zout)
        for k in range(Nin + 1):
            assert np.allclose(I @ zin ** k, zout ** k, atol=1e-13)


def test_kershaw_map_properties(orc):
    # eps = 1 is the identity; any eps maps the unit square faces onto themselves
    rng = np.random.default_rng(1)
    for _ in range(200):
        x, y, z = rng.uniform(0, 1, 3)
        assert np.allclose(orc.kershaw_map(1.0, x, y, z), (x, y, z), atol=1e-15)
    for eps in [0.3, 0.05]:
        for x in np.linspace(0, 1, 13):
            for y in [0.0, 1.0]:
                xx, yy, zz = orc.kershaw_map(eps, x, y, y)
                assert xx == x and abs(yy - y) < 1e-15 and abs(zz - y) < 1e-15
        # monotone in y for fixed x
        for x in np.linspace(0, 0.999, 7):
            ys = [orc.kershaw_map(eps, x, y, 0.5)[1] for y in np.linspace(0, 1, 41)]
            assert np.all(np.diff(ys) > 0)


def test_geofacs_affine_closed_form(orc):
    nx, ny, nz, N = 2, 3, 4, 4
    m = orc.BoxMesh(nx, ny, nz, N, "none")
    hx, hy, hz = 1.0 / nx, 1.0 / ny, 1.0 / nz
    G = m.ggeo.reshape(m.E, 7, m.Np)
    w = m.w
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    J = hx * hy * hz / 8.0
    JW = J * W
    assert np.allclose(G[:, orc.GWJ], JW, rtol=1e-13)
    assert np.allclose(G[:, orc.G00], JW * (2 / hx) ** 2, rtol=1e-13)
    assert np.allclose(G[:, orc.G11], JW * (2 / hy) ** 2, rtol=1e-13)
    assert np.allclose(G[:, orc.G22], JW * (2 / hz) ** 2, rtol=1e-13)
    for gid in (orc.G01, orc.G02, orc.G12):
        assert np.max(np.abs(G[:, gid])) < 1e-12


@pytest.mark.parametrize("mp", ["kershaw", "sine", "none"])
def test_volume_is_one(orc, mp):
    m = orc.BoxMesh(3, 3, 3, 5, mp, 0.3)
    vol = m.ggeo.reshape(m.E, 7, m.Np)[:, orc.GWJ].sum()
    tol = 1e-13 if mp != "sine" else 1e-6   # sine map: curved, quadrature error only
    assert abs(vol - 1.0) < tol


def _linear_field(m, a, b=0.3):
    return a[0] * m.x + a[1] * m.y + a[2] * m.z + b


@pytest.mark.parametrize("N", [2, 3, 5, 7])
def test_ax_energy_of_linear_field(orc, N):
    # u = a.x + b: sum_e u_e^T A_e u_e = |a|^2 vol exactly on trilinear elements (N >= 2)
    m = orc.BoxMesh(3, 2, 2, N, "kershaw", 0.3)
    a = np.array([0.7, -1.3, 0.4])
    u = _linear_field(m, a)
    Au = orc.ax(N, m.ggeo, m.D, u)
    assert abs(np.dot(u, Au) - np.dot(a, a)) < 1e-12


@pytest.mark.parametrize("N", [1, 3, 7])
def test_ax_constants_in_kernel_and_symmetry(orc, N):
    m = orc.BoxMesh(2, 2, 2, N, "kershaw", 0.3)
    one = np.ones(m.Nlocal)
    assert np.max(np.abs(orc.ax(N, m.ggeo, m.D, one))) < 1e-11
    rng = np.random.default_rng(N)
    p, q = rng.uniform(-1, 1, (2, m.Nlocal))
    Ap, Aq = orc.ax(N, m.ggeo, m.D, p), orc.ax(N, m.ggeo, m.D, q)
    assert abs(np.dot(q, Ap) - np.dot(p, Aq)) < 1e-12 * np.abs(np.dot(q, Ap)) + 1e-12
    # Helmholtz with lambda0 = 0, lambda1 = 1 is the lumped mass: A 1 = GwJ
    M1 = orc.ax(N, m.ggeo, m.D, one, lambda0=0.0, lambda1=1.0, poisson=False)
    assert np.allclose(M1, m.ggeo.reshape(m.E, 7, m.Np)[:, orc.GWJ].ravel(), rtol=1e-14)


def test_build_diag_brute_force(orc):
    N = 2
    m = orc.BoxMesh(2, 1, 1, N, "kershaw", 0.3)
    d = orc.build_diag(N, m.ggeo, m.D, lambda0=1.7, lambda1=0.4, poisson=False)
    diag = np.zeros(m.Nlocal)
    for n in range(m.Nlocal):
        e = np.zeros(m.Nlocal)
        e[n] = 1.0
        diag[n] = orc.ax(N, m.ggeo, m.D, e, lambda0=1.7, lambda1=0.4, poisson=False)[n]
    assert np.allclose(d, diag, rtol=1e-13)


@pytest.mark.parametrize("poisson", [True, False])
def test_build_diag_fields_brute_force(orc, poisson):
    """Per-node coefficients (ELLIPTIC COEFF FIELD): the diagonal equals
    e_n . A e_n with the field operator; the field form with constant fields
    equals the constant form."""
    N = 3
    m = orc.BoxMesh(2, 1, 1, N, "kershaw", 0.3)
    rng = np.random.default_rng(7)
    l0 = rng.uniform(0.5, 2.0, m.Nlocal)
    l1 = rng.uniform(0.0, 3.0, m.Nlocal)
    d = orc.build_diag(N, m.ggeo, m.D, l0, l1, poisson=poisson)
    diag = np.zeros(m.Nlocal)
    for n in range(m.Nlocal):
        e = np.zeros(m.Nlocal)
        e[n] = 1.0
        diag[n] = orc.ax(N, m.ggeo, m.D, e, l0, l1, poisson=poisson)[n]
    assert np.allclose(d, diag, rtol=1e-13)
    c0, c1 = np.full(m.Nlocal, 1.7), np.full(m.Nlocal, 0.4)
    assert np.allclose(orc.build_diag(N, m.ggeo, m.D, c0, c1, poisson=poisson),
                       orc.build_diag(N, m.ggeo, m.D, 1.7, 0.4, poisson=poisson), rtol=1e-14)


def test_ogs_rows_and_brute_force(orc):
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 30, 200).astype(np.int64)
    og = orc.Ogs(ids)
    # rows ordered by first appearance, members ascending, zeros excluded
    firsts = [og.ids[og.offsets[r]] for r in range(og.Nrows)]
    assert firsts == sorted(firsts)
    for r in range(og.Nrows):
        mem = og.ids[og.offsets[r]:og.offsets[r + 1]]
        assert np.all(np.diff(mem) > 0) and len(set(ids[mem])) == 1 and ids[mem[0]] != 0
    assert og.Nids == np.count_nonzero(ids)
    v = rng.uniform(-1, 1, 200)
    out = og.gs(v, "add")
    for i in range(200):
        if ids[i] == 0:
            assert out[i] == v[i]
        else:
            assert abs(out[i] - v[ids == ids[i]].sum()) < 1e-14
    mult = og.gs(np.ones(200), "add")
    assert np.allclose(og.invDegree[ids != 0] * mult[ids != 0], 1.0)
    assert np.all(og.invDegree[ids == 0] == 1.0)


def test_gs_sanity_box(orc):
    # src/mesh/meshParallelGatherScatterSetup.cpp:136-165: sum(invDegree * gs(1)) = E*Np
    m = orc.BoxMesh(3, 2, 4, 3, "kershaw", 0.3)
    og = orc.Ogs(m.gid)
    s = np.sum(og.invDegree * og.gs(np.ones(m.Nlocal), "add"))
    assert abs(s - m.Nlocal) < 1e-15 * m.Nlocal
    # number of unique nodes of the structured grid
    assert og.Nrows == (3 * 3 + 1) * (2 * 3 + 1) * (4 * 3 + 1)


def test_mask_is_boundary(orc):
    nx, ny, nz, N = 3, 2, 2, 3
    m = orc.BoxMesh(nx, ny, nz, N, "kershaw", 0.3)
    og = orc.Ogs(m.gid)
    mids = orc.mask_ids(m, og)
    g = m.gid - 1
    gx, gy = nx * N + 1, ny * N + 1
    ix, iy, iz = g % gx, (g // gx) % gy, g // (gx * gy)
    onb = (ix == 0) | (ix == nx * N) | (iy == 0) | (iy == ny * N) | (iz == 0) | (iz == nz * N)
    assert np.array_equal(np.sort(mids), np.nonzero(onb)[0])


@pytest.mark.parametrize("Nf,Nc", [(7, 3), (3, 1), (7, 5), (9, 5)])
def test_transfer_exact_and_adjoint(orc, Nf, Nc):
    m_f = orc.BoxMesh(2, 1, 1, Nf, "none")
    m_c = orc.BoxMesh(2, 1, 1, Nc, "none")
    R = orc.transfer_R(Nf, Nc)
    # prolongation of a degree-Nc polynomial is exact
    f = lambda x, y, z: (x ** Nc) - 2 * y * z + 0.5 * (z ** max(Nc - 1, 0))  # noqa: E731
    qc = f(m_c.x, m_c.y, m_c.z)
    qf = orc.prolongate(Nf, Nc, R, qc, np.zeros(m_f.Nlocal))
    assert np.allclose(qf, f(m_f.x, m_f.y, m_f.z), atol=1e-12)
    # coarsen is the transpose of prolongate
    rng = np.random.default_rng(0)
    a, b = rng.uniform(-1, 1, m_f.Nlocal), rng.uniform(-1, 1, m_c.Nlocal)
    lhs = np.dot(orc.coarsen(Nf, Nc, R, a), b)
    rhs = np.dot(a, orc.prolongate(Nf, Nc, R, b, np.zeros(m_f.Nlocal)))
    assert abs(lhs - rhs) < 1e-12 * max(1, abs(lhs))


def _assembled(el):
    """Dense global matrix of the masked operator on unique unmasked nodes."""
    m = el.mesh
    keep = np.ones(m.Nlocal, bool)
    keep[el.maskIds] = False
    uniq = np.unique(m.gid[keep])
    col = {g: c for c, g in enumerate(uniq)}
    n = len(uniq)
    A = np.zeros((n, n))
    for c, g in enumerate(uniq):
        e = ((m.gid == g) & keep).astype(float)
        Ae = el.operator(e)
        for i in np.nonzero(keep)[0]:
            A[col[m.gid[i]], c] = Ae[i]
    return A, uniq, keep


def test_operator_assembled_is_spd(orc):
    m = orc.BoxMesh(2, 2, 2, 2, "kershaw", 0.3)
    el = orc.Elliptic(m)
    A, _, _ = _assembled(el)
    assert np.allclose(A, A.T, atol=1e-12)
    assert np.linalg.eigvalsh(A).min() > 0


@pytest.mark.parametrize("precon", ["jacobi", "none"])
def test_pcg_matches_textbook_cg(orc, precon):
    m = orc.BoxMesh(2, 2, 2, 2, "kershaw", 0.3)
    el = orc.Elliptic(m)
    A, uniq, keep = _assembled(el)
    rng = np.random.default_rng(5)
    rhs = rng.uniform(-1, 1, m.Nlocal)
    x, info = el.solve(rhs, precon=precon, tol=1e-10, maxit=200)
    # textbook PCG on the assembled system with b = Q^T (masked rhs)
    b = np.array([rhs[(m.gid == g) & keep].sum() for g in uniq])
    Minv = 1.0 / np.diag(A) if precon == "jacobi" else np.ones(len(b))
    xs = np.zeros_like(b); r = b.copy(); z = Minv * r; p = z.copy(); rz = r @ z
    hist = []
    for it in range(200):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        xs += alpha * p
        r -= alpha * Ap
        hist.append(np.sqrt(r @ r / el.volume))
        if hist[-1] <= 1e-10 * np.sqrt(b @ b / el.volume):
            break
        z = Minv * r
        rz, rz_old = r @ z, rz
        if precon == "none":
            break   # the reference's unpreconditioned variant is not textbook CG; see below
        p = z + (rz / rz_old) * p
    if precon == "jacobi":
        assert abs(info["iterations"] - len(hist)) <= 0
        assert np.allclose(info["history"][:5], hist[:5], rtol=1e-10)
        xg = np.array([x[(m.gid == g) & keep][0] for g in uniq])
        assert np.allclose(xg, np.linalg.solve(A, b), rtol=1e-8, atol=1e-10)
    else:
        # PCG.cpp:128-130 sets rdotz1 to the normalised residual norm: first
        # step length is ||r||_n / p^T A p instead of r^T r / p^T A p
        r0 = np.sqrt(b @ b / el.volume)
        alpha = r0 / (b @ (A @ b))
        r1 = b - alpha * (A @ b)
        assert abs(info["history"][0] - np.sqrt(r1 @ r1 / el.volume)) < 1e-12 * r0


def test_manufactured_solution_converges(orc):
    errs = []
    for N in [3, 5, 7]:
        m = orc.BoxMesh(3, 3, 3, N, "sine", 0.0)
        el = orc.Elliptic(m)
        gw = m.ggeo.reshape(m.E, 7, m.Np)[:, orc.GWJ].ravel()
        u = np.sin(np.pi * m.x) * np.sin(np.pi * m.y) * np.sin(np.pi * m.z)
        rhs = gw * 3 * np.pi ** 2 * u
        x, info = el.solve(rhs, tol=1e-12, maxit=2000)
        errs.append(np.max(np.abs(x - u)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-5


def test_field_elliptic_constant_fields_equal_constant_operator(orc):
    """The oracle's field operator (ax with fields, mask, gather-scatter) with
    constant fields equals its constant-coefficient operator (orc_operator)."""
    m = orc.BoxMesh(3, 2, 2, 4, "kershaw", 0.3)
    q = np.random.default_rng(3).uniform(-1, 1, m.Nlocal)
    ref = orc.Elliptic(m, 1.7, 0.4, False)
    fe = orc.FieldElliptic(m, np.full(m.Nlocal, 1.7), np.full(m.Nlocal, 0.4), poisson=False)
    a, b = fe.operator(q), ref.operator(q)
    assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    assert np.allclose(fe.inv_diag(), ref.inv_diag(), rtol=1e-14)
