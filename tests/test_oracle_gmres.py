"""Pinning of the PGMRES oracle (oracle.mg.gmres_solve) by what the
mathematics fixes: it reaches the solution of the assembled system; its
Givens residual estimate equals the true residual; it minimises the residual
over the Krylov space, so after k steps its residual is no larger than CG's
(same inner product, no preconditioner); restarted runs still converge and
the non-flexible and flexible forms agree for a fixed linear preconditioner."""
import numpy as np


def _setup(orc):
    mo = orc.BoxMesh(3, 3, 2, 4, "kershaw", 0.3)
    eo = orc.Elliptic(mo, 1.0, 2.0, False)
    rhs = np.random.default_rng(0).uniform(-1, 1, mo.Nlocal)
    return mo, eo, rhs


def test_gmres_reaches_cg_solution(orc):
    import oracle.mg as omg
    mo, eo, rhs = _setup(orc)
    xc, _ = omg.fcg_solve(eo, rhs, lambda r: eo.inv_diag() * r, flexible=False, tol=1e-13,
                          maxit=5000)
    for flexible in (True, False):
        xg, info = omg.gmres_solve(eo, rhs, lambda r: eo.inv_diag() * r, flexible=flexible,
                                   restart=30, tol=1e-12, maxit=3000)
        assert np.max(np.abs(xg - xc)) <= 1e-9 * np.max(np.abs(xc))


def test_gmres_estimate_equals_true_residual(orc):
    import oracle.mg as omg
    mo, eo, rhs = _setup(orc)
    x, info = omg.gmres_solve(eo, rhs, lambda r: r, flexible=True, restart=40, tol=1e-6, maxit=40)
    mask = np.isin(np.arange(mo.Nlocal), eo.maskIds)
    b = eo.ogs.gs(np.where(mask, 0.0, rhs), "add")
    r = b - eo.operator(x)
    true = np.sqrt(np.sum(eo.invDegree * r * r) / eo.volume)
    assert abs(true - info["res"]) <= 1e-8 * info["res0"]


def test_gmres_minimal_residual_beats_cg(orc):
    import oracle.mg as omg
    mo, eo, rhs = _setup(orc)
    _, cg = omg.fcg_solve(eo, rhs, lambda r: r, flexible=False, tol=1e-30, maxit=12)
    for k in range(1, 13):
        _, g = omg.gmres_solve(eo, rhs, lambda r: r, flexible=True, restart=30, tol=1e-30, maxit=k)
        assert g["res"] <= cg["history"][k - 1] * (1 + 1e-10)


def test_restarted_gmres_converges(orc):
    import oracle.mg as omg
    mo, eo, rhs = _setup(orc)
    x5, i5 = omg.gmres_solve(eo, rhs, lambda r: eo.inv_diag() * r, flexible=False, restart=5,
                             tol=1e-10, maxit=5000)
    x30, i30 = omg.gmres_solve(eo, rhs, lambda r: eo.inv_diag() * r, flexible=False, restart=30,
                               tol=1e-10, maxit=5000)
    assert i5["iterations"] >= i30["iterations"]
    assert np.max(np.abs(x5 - x30)) <= 1e-7 * np.max(np.abs(x30))
