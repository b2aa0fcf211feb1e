"""Pinning of the multigrid oracle (oracle/mg.py) by what the mathematics fixes:
the Chebyshev smoother's error propagator equals the scaled Chebyshev
polynomial of S A, the Arnoldi bound brackets the dense spectrum, the V-cycle
with an exact coarse solve is a symmetric operator in the assembled inner
product, and MG-preconditioned flexible CG reaches the same solution as the
independently pinned Jacobi PCG."""
import numpy as np
import pytest


def _dense(op, gid, keep):
    """Matrix of op on unique unmasked dofs (inputs continuous: 1 at every copy)."""
    cols = []
    for c in np.nonzero(keep)[0]:
        e = (gid == gid[c]).astype(float)
        cols.append(op(e)[keep])
    return np.array(cols).T


@pytest.fixture(scope="module")
def small(orc):
    import oracle.mg as omg
    mo = orc.BoxMesh(2, 2, 2, 3, "kershaw", 0.3)
    om = omg.MG(mo, [3, 1], T=np.float64, coarse_tol=1e-14, coarse_maxit=1000)
    return mo, om


def test_arnoldi_bound_brackets_spectrum(small):
    mo, om = small
    L = om.levels[0]
    # unique unmasked nodes: one representative per gather row
    keep = np.zeros(mo.Nlocal, bool)
    seen = set()
    masked = set(L.el.maskIds.tolist())
    for i, g in enumerate(mo.gid):
        if i not in masked and g not in seen:
            seen.add(g)
            keep[i] = True
    SA = _dense(lambda v: L.invDiagT * om.op(L, v), mo.gid, keep)
    ev = np.linalg.eigvals(SA).real
    # Ritz values of the Arnoldi process (standard inner product on unique
    # dofs, where S A is not normal) lie in the field of values of S A
    fov_max = np.linalg.eigvalsh(0.5 * (SA + SA.T)).max()
    assert L.rho <= fov_max * (1 + 1e-10)
    assert L.rho >= 0.95 * ev.max()


def test_chebyshev_error_propagator(small):
    mo, om = small
    L = om.levels[0]
    keep = np.zeros(mo.Nlocal, bool)
    seen = set()
    masked = set(L.el.maskIds.tolist())
    for i, g in enumerate(mo.gid):
        if i not in masked and g not in seen:
            seen.add(g)
            keep[i] = True
    A = _dense(lambda v: om.op(L, v), mo.gid, keep)
    S = np.diag(L.invDiagT[keep])
    # smoother with x = 0: x = q(SA) S b; error propagator E = I - q(SA) S A
    col = {g: c for c, g in enumerate(mo.gid[keep])}
    idx = np.array([col.get(g, -1) for g in mo.gid])
    def smooth_unique(b_u):
        b = np.where(idx >= 0, b_u[np.maximum(idx, 0)], 0.0)   # continuous representation
        b[L.el.maskIds] = 0
        x = np.zeros(mo.Nlocal)
        om.smooth(L, b, x, True)
        return x[keep]
    n = keep.sum()
    E = np.eye(n) - np.array([smooth_unique(c) for c in A.T]).T
    theta, delta = 0.5 * (L.lmax + L.lmin), 0.5 * (L.lmax - L.lmin)
    lam, V = np.linalg.eig(S @ A)
    k = L.degree
    T = np.polynomial.chebyshev.Chebyshev.basis(k)
    expect = T((theta - lam.real) / delta) / T(theta / delta)
    got = np.diag(np.linalg.solve(V, E @ V)).real
    assert np.allclose(got, expect, atol=1e-8)


def test_vcycle_symmetric_with_exact_coarse(small, orc):
    mo, om = small
    el = om.levels[0].el
    rng = np.random.default_rng(0)
    def assembled():
        r = rng.uniform(-1, 1, mo.Nlocal)
        r[el.maskIds] = 0
        return el.ogs.gs(r, "add")
    r1, r2 = assembled(), assembled()
    z1, z2 = om.apply(r1), om.apply(r2)
    w = el.invDegree
    a, b = np.sum(w * z1 * r2), np.sum(w * r1 * z2)
    assert abs(a - b) < 1e-10 * abs(a)
    # and positive definite on these samples
    assert np.sum(w * z1 * r1) > 0 and np.sum(w * z2 * r2) > 0


def test_mg_fcg_matches_jacobi_solution(orc):
    import oracle.mg as omg
    mo = orc.BoxMesh(3, 3, 3, 7, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    rhs = np.random.default_rng(7).uniform(-1, 1, mo.Nlocal)
    xj, ij = eo.solve(rhs, tol=1e-12, maxit=5000)
    for levels in ([7, 3, 1], [7, 5, 3, 1]):
        om = omg.MG(mo, levels)
        xm, im = omg.fcg_solve(eo, rhs, om.apply, tol=1e-12, maxit=400)
        assert im["iterations"] < ij["iterations"] / 4
        assert np.max(np.abs(xj - xm)) <= 1e-9 * np.max(np.abs(xj))


def test_splitmix_reference_values():
    import oracle.mg as omg
    # splitmix64 with state 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4
    v = omg.splitmix_uniform(0, 2)
    assert v[0] == (0xE220A8397B1DCDAF >> 11) / 2.0 ** 53
    assert v[1] == (0x6E789E6AA1B965F4 >> 11) / 2.0 ** 53


@pytest.mark.parametrize("shape,unknowns", [((2, 2, 2), 1), ((3, 2, 2), 2), ((1, 1, 1), 0)])
def test_arnoldi_breakdown_on_tiny_levels(orc, shape, unknowns):
    """A smoothed coarse level with fewer unknowns than Arnoldi steps: the
    Krylov space becomes invariant, and the leading block's eigenvalues are the
    exact spectrum of S A (1 for Jacobi on one unknown); a level with no
    unknowns gets rho = 1 and smooths to zero."""
    import oracle.mg as omg
    mo = orc.BoxMesh(*shape, 3, "kershaw", 0.2)
    om = omg.MG(mo, [3, 1], T=np.float64, coarse="smooth", lambda0=1.0, lambda1=2.5,
                poisson=False)
    L = om.levels[-1]
    keep = np.zeros(L.E * L.Np, bool)
    seen = set()
    masked = set(L.el.maskIds.tolist())
    for i, g in enumerate(L.mesh.gid):
        if i not in masked and g not in seen:
            seen.add(g)
            keep[i] = True
    assert keep.sum() == unknowns
    if unknowns == 0:
        assert L.rho == 1.0
    else:
        SA = _dense(lambda v: L.invDiagT * om.op(L, v), L.mesh.gid, keep)
        exact = np.abs(np.linalg.eigvals(SA)).max()
        assert abs(L.rho - exact) <= 1e-10 * exact
        if unknowns == 1:
            assert abs(L.rho - 1.0) <= 1e-12
    rhs = np.random.default_rng(1).uniform(-1, 1, mo.Nlocal)
    eo = orc.Elliptic(mo, 1.0, 2.5, False)
    z = om.apply(eo.ogs.gs(np.where(np.isin(np.arange(mo.Nlocal), eo.maskIds), 0.0, rhs), "add"))
    assert np.all(np.isfinite(z))


def test_helmholtz_schwarz_coarse_rho_mesh_independent(orc):
    """The Helmholtz box solve drops the constant mode of a box with free ends
    (as the Poisson one does): on undeformed meshes the order-1 level's
    rho(S A) then stays bounded under refinement instead of growing like
    1/(lambda1 h^2); with lambda1 = 0 it equals the Poisson smoother."""
    import oracle.mg as omg
    rhos = []
    for n in (4, 8):
        mo = orc.BoxMesh(n, n, n, 3, "kershaw", 1.0)
        om = omg.MG(mo, [3, 1], T=np.float64, degree=3, cheby="opt4", S="ras", coarse="smooth",
                    lambda0=1.0, lambda1=4.0, poisson=False)
        rhos.append(om.levels[-1].rho)
    assert max(rhos) < 2.5 and abs(rhos[1] - rhos[0]) < 0.2, rhos
    mo = orc.BoxMesh(4, 4, 4, 3, "kershaw", 0.5)
    a = omg.MG(mo, [3, 1], T=np.float64, degree=3, cheby="opt4", S="asm", coarse="smooth")
    b = omg.MG(mo, [3, 1], T=np.float64, degree=3, cheby="opt4", S="asm", coarse="smooth",
               lambda0=1.0, lambda1=0.0, poisson=False)
    for La, Lb in zip(a.levels, b.levels):
        assert np.array_equal(La.schwarz.invL, Lb.schwarz.invL)


def test_jacobi_smoothed_coarse_level(orc):
    """coarse="smooth_jacobi": Schwarz on the finer levels, Chebyshev-Jacobi on
    the order-1 level (whose rho is then that of D^-1 A, below 2), and the
    preconditioned solve converges to the Jacobi-PCG solution."""
    import oracle.mg as omg
    mo = orc.BoxMesh(6, 5, 5, 3, "kershaw", 0.1)
    om = omg.MG(mo, [3, 1], T=np.float32, degree=3, cheby="opt4", S="ras",
                coarse="smooth_jacobi", lambda0=1.0, lambda1=4.0, poisson=False)
    Lc = om.levels[-1]
    assert getattr(Lc, "schwarz", None) is None and Lc.rho < 2.0
    assert om.levels[0].schwarz is not None
    eo = orc.Elliptic(mo, 1.0, 4.0, False)
    rhs = np.random.default_rng(2).uniform(-1, 1, mo.Nlocal)
    x, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=300)
    xj, ij = omg.fcg_solve(eo, rhs, lambda r: eo.inv_diag() * r, flexible=False, tol=1e-12,
                           maxit=5000)
    assert info["iterations"] < ij["iterations"] // 4
    assert np.max(np.abs(x - xj)) <= 1e-7 * np.max(np.abs(xj))


def _unique_keep(L):
    keep = np.zeros(L.E * L.Np, bool)
    seen = set()
    masked = set(L.el.maskIds.tolist())
    for i, g in enumerate(L.mesh.gid):
        if i not in masked and g not in seen:
            seen.add(g)
            keep[i] = True
    return keep


def _fourth_kind_E(k, lam, rho):
    """E_k(lam) = W_k(1 - 2 lam / rho) / (2k + 1), W_k(cos t) = sin((k + 1/2) t) / sin(t / 2),
    E_0 = 1: the error propagator of the (unoptimised) 4th-kind Chebyshev
    smoother of degree k (Lottes, arXiv:2202.08830, eq. for the 4th-kind
    polynomials); evaluated through the three-term recurrence of W_k,
    W_0 = 1, W_1 = 2x + 1, W_{j+1} = 2x W_j - W_{j-1}, which is valid for every real x."""
    x = 1.0 - 2.0 * lam / rho
    w0, w1 = np.ones_like(x), 2.0 * x + 1.0
    if k == 0:
        return w0
    for _ in range(k - 1):
        w0, w1 = w1, 2.0 * x * w1 - w0
    return w1 / (2.0 * k + 1.0)


@pytest.mark.parametrize("degree", [2, 3, 4])
def test_fourth_kind_chebyshev_error_propagator(orc, monkeypatch, degree):
    """The optimal 4th-kind smoother (smoothFourthKindChebyshev,
    ellipticMultiGridLevel.cpp:212-270) pinned by its closed form, on the
    eigenbasis of S A: with every beta = 1 the error propagator is E_k(lambda);
    with a one-hot beta = e_j it is 1 - E_{j-1}(lambda) + E_j(lambda) (the j-th
    update direction alone).  A shifted beta index, a wrong rC or dC, or a
    wrong 4/(3 rho) start changes these polynomials and fails."""
    import oracle.mg as omg
    mo = orc.BoxMesh(2, 2, 2, 3, "kershaw", 0.3)
    om = omg.MG(mo, [3, 1], T=np.float64, degree=degree, cheby="opt4", coarse_tol=1e-14,
                coarse_maxit=1000)
    L = om.levels[0]
    keep = _unique_keep(L)
    A = _dense(lambda v: om.op(L, v), mo.gid, keep)
    S = np.diag(L.invDiagT[keep])
    col = {g: c for c, g in enumerate(mo.gid[keep])}
    idx = np.array([col.get(g, -1) for g in mo.gid])

    def propagator():
        def smooth_unique(b_u):
            b = np.where(idx >= 0, b_u[np.maximum(idx, 0)], 0.0)
            b[L.el.maskIds] = 0
            x = np.zeros(mo.Nlocal)
            om.smooth(L, b, x, True)
            return x[keep]
        n = keep.sum()
        return np.eye(n) - np.array([smooth_unique(c) for c in A.T]).T

    lam, V = np.linalg.eig(S @ A)
    lam = lam.real
    rho = L.lmax
    assert lam.max() < rho and lam.min() > 0
    cases = [(np.ones(degree), _fourth_kind_E(degree, lam, rho))]
    for j in range(1, degree + 1):
        beta = np.zeros(degree)
        beta[j - 1] = 1.0
        cases.append((beta, 1.0 - _fourth_kind_E(j - 1, lam, rho) + _fourth_kind_E(j, lam, rho)))
    for beta, expect in cases:
        monkeypatch.setitem(omg.OPT4_BETAS, degree, list(beta))
        E = propagator()
        got = np.diag(np.linalg.solve(V, E @ V)).real
        assert np.allclose(got, expect, atol=1e-9), (beta, np.max(np.abs(got - expect)))
        # and E is diagonal in that basis (a polynomial in S A)
        off = np.linalg.solve(V, E @ V)
        assert np.max(np.abs(off - np.diag(np.diag(off)))) < 1e-8
    monkeypatch.undo()
    # the optimised betas: the propagator stays a polynomial in S A with |E| < 1 on (0, rho]
    E = propagator()
    got = np.diag(np.linalg.solve(V, E @ V)).real
    assert np.max(np.abs(got)) < 1.0
    expect = np.zeros_like(lam)
    b = omg.OPT4_BETAS[degree]
    for j in range(1, degree + 1):
        expect += b[j - 1] * (_fourth_kind_E(j - 1, lam, rho) - _fourth_kind_E(j, lam, rho))
    assert np.allclose(got, 1.0 - expect, atol=1e-9)


def test_fourth_kind_polynomials_closed_form():
    """The recurrence used above against the trigonometric definition."""
    t = np.linspace(0.1, 3.0, 17)
    rho = 2.0
    lam = 0.5 * rho * (1.0 - np.cos(t))          # 1 - 2 lam / rho = cos t
    for k in range(0, 7):
        w = np.sin((k + 0.5) * t) / np.sin(0.5 * t)
        assert np.allclose(_fourth_kind_E(k, lam, rho), w / (2 * k + 1), atol=1e-13)
