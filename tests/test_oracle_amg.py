"""Pinning of the coarse-level AMG oracle (oracle/amg.py) by what the
mathematics fixes: the assembled matrix equals the matrix-free order-1
operator (gather-scatter of the element operator, masked) on continuous
vectors; the hierarchy is Galerkin (symmetric coarse matrices, P of full
column rank with the aggregates partitioning the rows); the V-cycle is a
symmetric positive definite operator that contracts the error; PCG with it
reaches the direct solution; the multigrid solve with the AMG coarse level
converges like the one with the Jacobi-PCG coarse solve."""
import numpy as np
import pytest
import scipy.sparse.linalg as spl


@pytest.fixture(scope="module")
def level(orc):
    import oracle.mg as omg
    mo = orc.BoxMesh(6, 5, 4, 3, "kershaw", 0.3)
    om = omg.MG(mo, [3, 1], T=np.float64, coarse="amg", lambda0=1.0, lambda1=0.7, poisson=False)
    return mo, om, om.levels[-1]


def test_assembled_matrix_is_the_operator(level):
    mo, om, L = level
    A, rows = L.amg_A, L.amg_rows
    rng = np.random.default_rng(0)
    xu = rng.uniform(-1, 1, A.shape[0])
    xe = np.where(rows >= 0, xu[np.maximum(rows, 0)], 0.0)
    ye = om.op(L, xe)                    # masked, gathered-scattered operator (FP64 level)
    assert np.allclose(ye[L.amg_first], A @ xu, rtol=0, atol=1e-13 * np.abs(ye).max())
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()
    # every unmasked node is in exactly one row; masked nodes in none
    assert np.all(rows[L.el.maskIds] == -1)
    assert np.all(np.bincount(rows[rows >= 0]) >= 1)


def test_hierarchy_is_galerkin(level):
    import oracle.amg as oamg
    _, _, L = level
    H = oamg.AMG(L.amg_A, coarsest=8)
    assert len(H.levels) >= 3
    for k in range(len(H.levels) - 1):
        lv, nx = H.levels[k], H.levels[k + 1]
        P = lv["P"]
        assert P.shape == (lv["A"].shape[0], nx["A"].shape[0])
        Ac = (P.T @ lv["A"] @ P)
        assert abs(Ac - nx["A"]).max() <= 1e-13 * abs(nx["A"]).max()
        assert abs(nx["A"] - nx["A"].T).max() <= 1e-13 * abs(nx["A"]).max()
        assert 0.0 < lv["omega"] < 2.0
    S = __import__("oracle.amg", fromlist=["x"])
    agg, na = S.aggregate(S.strength(H.levels[0]["A"]))
    assert agg.min() == 0 and agg.max() == na - 1 and np.all(np.bincount(agg) >= 1)


def test_vcycle_spd_and_contracting(level):
    import oracle.amg as oamg
    _, _, L = level
    A = L.amg_A
    H = oamg.AMG(A, coarsest=8)
    rng = np.random.default_rng(1)
    u, v = rng.standard_normal(A.shape[0]), rng.standard_normal(A.shape[0])
    assert abs(u @ H.vcycle(v) - v @ H.vcycle(u)) <= 1e-11 * np.linalg.norm(u) * np.linalg.norm(v)
    assert u @ H.vcycle(u) > 0
    # stationary iteration x += M (b - A x): error contracts in the A norm
    xs = spl.spsolve(A.tocsc(), v)
    x = np.zeros_like(v)
    e0 = np.sqrt((xs - x) @ (A @ (xs - x)))
    for _ in range(6):
        x = x + H.vcycle(v - A @ x)
    e6 = np.sqrt((xs - x) @ (A @ (xs - x)))
    assert e6 < 0.8 ** 6 * e0


def test_pcg_amg_reaches_direct_solution(level, orc):
    mo, om, L = level
    A = L.amg_A
    b = np.random.default_rng(2).uniform(-1, 1, A.shape[0])
    xs = spl.spsolve(A.tocsc(), b)
    om.coarse_tol, om.coarse_maxit = 1e-12, 100
    be = np.zeros(L.E * L.Np)
    be[L.amg_rows >= 0] = b[L.amg_rows[L.amg_rows >= 0]]
    xe, it = om.coarse_solve(L, be)
    assert it < 30
    assert np.max(np.abs(xe[L.amg_first] - xs)) <= 1e-10 * np.max(np.abs(xs))
    assert np.all(xe[L.el.maskIds] == 0)


def test_mg_with_amg_coarse_converges_like_pcg_coarse(orc):
    import oracle.mg as omg
    mo = orc.BoxMesh(6, 6, 5, 5, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    rhs = np.random.default_rng(3).uniform(-1, 1, mo.Nlocal)
    its = {}
    for co in ("pcg", "amg"):
        om = omg.MG(mo, [5, 3, 1], T=np.float32, degree=2, coarse=co)
        x, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
        its[co] = info["iterations"]
        if co == "pcg":
            xp = x
        else:
            assert np.max(np.abs(x - xp)) <= 1e-8 * np.max(np.abs(xp))
    assert abs(its["amg"] - its["pcg"]) <= 2, its
