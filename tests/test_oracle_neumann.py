"""Pinning of the oracle's all-Neumann path (ellipticSetup.cpp:187-206,
ellipticZeroMean.cpp:31-45 at ellipticSolve.cpp:101,188-189 and
ellipticPreconditioner.cpp:82-83) by what the mathematics fixes.

On a closed box (NEUMANN on every face) and on a box periodic in all three
directions the Poisson matrix is singular with the constant vector as null
space.  Removing the mean of the unassembled residual makes its assembled
form consistent (its sum over the unique nodes is the sum over the local
nodes, zero), so PCG converges; the returned x has zero E-vector mean and
solves A x = gs(b - mean(b)); for the cosine modes of the box (zero normal
derivative on the faces) the discrete solution approaches the exact one minus
its mean.  The multigrid-preconditioned solve (coarse right-hand side made
consistent, pseudo-inverse on the AMG's dense level) reaches the Jacobi-PCG
solution.  A dropped zero-mean, a mean taken over the unique instead of the
local nodes, or an inverse instead of a pseudo-inverse fails one of these."""
import numpy as np
import pytest


def closed_problem(orc, n, N, map="none", eps=0.0, periodic=""):
    bc = "neumann" if periodic != "xyz" else "dirichlet"
    m = orc.BoxMesh(n, n, n, N, map, eps, bc=bc, periodic=periodic)
    el = orc.Elliptic(m)
    return m, el


@pytest.mark.parametrize("periodic", ["", "xyz"])
def test_all_neumann_detected(orc, periodic):
    m, el = closed_problem(orc, 3, 3, periodic=periodic)
    assert el.all_neumann and el.zero_mean and len(el.maskIds) == 0
    d = orc.BoxMesh(3, 3, 3, 3)
    assert not orc.Elliptic(d).all_neumann
    assert not orc.Elliptic(m, 1.0, 2.0, poisson=False).all_neumann   # Helmholtz: regular
    if periodic:
        # ids wrap: the periodic box of n^3 elements has (n N)^3 unique nodes
        assert len(np.unique(m.gid)) == (3 * 3) ** 3


@pytest.mark.parametrize("periodic,N", [("", 7), ("xyz", 7), ("", 5)])
def test_closed_box_pcg_manufactured(orc, periodic, N):
    n = 3
    m, el = closed_problem(orc, n, N, periodic=periodic)
    k = 2 * np.pi if periodic else np.pi
    u = np.cos(k * m.x) * np.cos(k * m.y) * np.cos(k * m.z)
    gw = m.ggeo.reshape(m.E, 7, m.Np)[:, orc.GWJ].ravel()
    b = gw * 3 * k * k * u
    x, info = el.solve(b, tol=1e-13, maxit=5000)
    assert info["iterations"] > 0 and info["res"] <= 1e-13 * info["res0"] * (1 + 1e-9)
    # zero E-vector mean (ellipticZeroMean over the local nodes)
    assert abs(np.sum(x)) <= 1e-12 * np.sum(np.abs(x))
    # A x equals the assembled, mean-free right-hand side
    bb = el.ogs.gs(b - np.sum(b) / b.size, "add")
    assert np.max(np.abs(el.operator(x) - bb)) <= 1e-10 * np.max(np.abs(bb))
    # the exact solution up to a constant
    ue = u - np.sum(u) / u.size
    err = np.max(np.abs(x - ue)) / np.max(np.abs(ue))
    assert err < (2e-4 if N == 5 else 2e-6), err


def test_consistency_needs_the_local_mean(orc):
    """A mean over the unique nodes leaves the assembled residual
    inconsistent: CG then stalls far above the tolerance."""
    import oracle.mg as omg
    m, el = closed_problem(orc, 3, 5, "kershaw", 0.3)
    b = np.random.default_rng(0).uniform(-1, 1, m.Nlocal)
    w = el.invDegree
    wrong = el.ogs.gs(b - np.sum(w * b) / np.sum(w), "add")       # not consistent
    right = el.ogs.gs(b - np.sum(b) / b.size, "add")
    assert abs(np.sum(w * right)) <= 1e-12 * np.sum(np.abs(b))
    assert abs(np.sum(w * wrong)) > 1e-6
    inv = el.inv_diag()
    el.zero_mean = True
    _, ok = omg.pcg(el, right.copy(), lambda r: inv * r, flexible=False, tol=1e-11, maxit=3000)
    _, bad = omg.pcg(el, wrong.copy(), lambda r: inv * r, flexible=False, tol=1e-11, maxit=300)
    assert ok["res"] <= 1e-11 * ok["res0"] * (1 + 1e-9)
    assert bad["res"] > 1e-6 * bad["res0"]


def test_amg_pseudo_inverse_on_singular_matrix(orc):
    import oracle.amg as amg
    import oracle.mg as omg
    m, el = closed_problem(orc, 4, 3, "kershaw", 0.2)
    om = omg.MG(m, [3, 1], T=np.float64, coarse="amg", coarse_tol=1e-2)
    A = om.levels[-1].amg_A
    ones = np.ones(A.shape[0])
    assert np.max(np.abs(A @ ones)) <= 1e-12 * abs(A).max()
    h = amg.AMG(A, coarsest=40, null_space=True)
    assert len(h.levels) >= 2
    Ac = h.levels[-1]["A"].toarray()
    Pi = h.levels[-1]["inv"]
    # Moore-Penrose conditions on the (singular) coarsest matrix
    assert np.allclose(Ac @ Pi @ Ac, Ac, atol=1e-10 * abs(Ac).max())
    assert np.allclose(Pi @ Ac @ Pi, Pi, atol=1e-8 * abs(Pi).max())
    # the fine constant is carried down as the coarse null vector B
    B = h.levels[-1]["B"]
    assert np.max(np.abs(Ac @ B)) <= 1e-12 * abs(Ac).max() * np.max(B)
    assert np.max(np.abs(Pi @ B)) <= 1e-8 * abs(Pi).max() * np.max(B)
    for k in range(len(h.levels) - 1):
        P = h.levels[k]["P"]
        assert np.allclose(P @ h.levels[k + 1]["B"], h.levels[k]["B"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("periodic,S,coarse", [("", "jacobi", "amg"), ("", "jacobi", "pcg"),
                                               ("xyz", "ras", "amg"), ("", "asm", "smooth")])
def test_mg_solve_reaches_jacobi_solution(orc, periodic, S, coarse):
    import oracle.mg as omg
    N = 5
    m, el = closed_problem(orc, 4, N, "none" if periodic else "kershaw", 0.0 if periodic else 0.3,
                           periodic=periodic)
    b = np.random.default_rng(3).uniform(-1, 1, m.Nlocal)
    xj, ij = el.solve(b, tol=1e-12, maxit=8000)
    om = omg.MG(m, [N, 3, 1], T=np.float64, degree=3, cheby="opt4", S=S, coarse=coarse,
                coarse_tol=1e-3 if coarse != "amg" else 0.0, coarse_maxit=3)
    xm, im = omg.fcg_solve(el, b, om.apply, flexible=True, tol=1e-12, maxit=300)
    assert im["iterations"] < ij["iterations"] / 4, (im["iterations"], ij["iterations"])
    assert abs(np.sum(xm)) <= 1e-12 * np.sum(np.abs(xm))
    assert np.max(np.abs(xm - xj)) <= 1e-9 * np.max(np.abs(xj))
