"""Pinning of the projection oracle (oracle.projection) by what the
mathematics fixes: the basis is A-orthonormal; the projected residual is
orthogonal to the basis; xbar equals the Galerkin solution over the span of
the raw corrections, computed densely without the Gram-Schmidt; a right-hand
side already solved leaves only the previous solve's residual; CLASSIC and
ACONJ give the same projection; the basis-size schedule and the discard of a
dependent vector follow SolutionProjection (ellipticSolutionProjection.cpp:
194-235, 269-288)."""
import numpy as np
import pytest


def _setup(orc, lam1=2.0):
    mo = orc.BoxMesh(3, 3, 2, 4, "kershaw", 0.3)
    eo = orc.Elliptic(mo, 1.0, lam1, False)
    return mo, eo


def _rhs(mo, t):
    return (np.sin(np.pi * mo.x + t) * np.cos(np.pi * mo.y - 0.5 * t) * (1 + mo.z * t)
            + 0.01 * t * t * mo.x * mo.y)


def _inner(orc, eo, tol=1e-12):
    import oracle.mg as omg
    inv = eo.inv_diag()
    return lambda r: omg.pcg(eo, r, lambda v: inv * v, False, tol, 5000)


def _A(eo):
    return eo.operator


@pytest.mark.parametrize("kind", [1, 2])
def test_basis_is_a_orthonormal_and_residual_orthogonal(orc, kind):
    import oracle.projection as op
    mo, eo = _setup(orc)
    pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=kind, max_vecs=6, start=1)
    inner = _inner(orc, eo)
    for k in range(5):
        rhs = _rhs(mo, 0.3 * k)
        r = op.residual(eo, rhs, np.zeros(mo.Nlocal))
        rp, ran = pj.pre(r.copy())
        if ran:
            for v in pj.xx:                       # Galerkin: <xx_k, r - A xbar> = 0
                assert abs(pj.dot(v, rp)) <= 1e-10 * np.sqrt(pj.dot(r, r) * pj.dot(v, v))
        dx, _ = inner(rp)
        pj.post(dx)
        G = np.array([[pj.dot(a, eo.operator(b)) for b in pj.xx] for a in pj.xx])
        assert np.max(np.abs(G - np.eye(len(pj.xx)))) <= 1e-9


@pytest.mark.parametrize("kind", [1, 2])
def test_xbar_is_galerkin_solution_over_raw_corrections(orc, kind):
    import oracle.projection as op
    mo, eo = _setup(orc)
    pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=kind, max_vecs=8, start=1)
    inner = _inner(orc, eo)
    raw = []
    for k in range(5):
        r = op.residual(eo, _rhs(mo, 0.25 * k), np.zeros(mo.Nlocal))
        rp, ran = pj.pre(r.copy())
        if ran:
            Q = np.array(raw).T                   # raw corrections (x + xbar span the same space)
            w = eo.invDegree
            AQ = np.array([eo.operator(q) for q in raw]).T
            c = np.linalg.solve(Q.T @ (w[:, None] * AQ), Q.T @ (w * r))
            y = Q @ c
            assert np.max(np.abs(pj.xbar - y)) <= 1e-8 * np.max(np.abs(y))
        dx, _ = inner(rp)
        full = pj.post(dx)
        raw.append(full.copy() if len(pj.xx) == 1 else dx.copy())


def test_repeated_rhs_leaves_previous_residual(orc):
    import oracle.projection as op
    mo, eo = _setup(orc)
    pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=2, max_vecs=8, start=1)
    inner = _inner(orc, eo, tol=1e-10)
    rhs = _rhs(mo, 0.7)
    r = op.residual(eo, rhs, np.zeros(mo.Nlocal))
    pj.pre(r.copy())
    dx, st = inner(r.copy())
    x1 = pj.post(dx)
    r1 = r - eo.operator(x1)                      # the first solve's final residual
    rp, ran = pj.pre(r.copy())
    assert ran and np.max(np.abs(rp - r1)) <= 1e-6 * np.max(np.abs(r1)) + 1e-14 * np.max(np.abs(r))
    assert np.max(np.abs(pj.xbar - x1)) <= 1e-12 * np.max(np.abs(x1))


def test_classic_equals_aconj(orc):
    import oracle.projection as op
    mo, eo = _setup(orc)
    inner = _inner(orc, eo)
    outs = {}
    for kind in (1, 2):
        pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=kind, max_vecs=3, start=2)
        xs = []
        for k in range(7):
            x, st = op.solve(eo, pj, _rhs(mo, 0.2 * k), np.zeros(mo.Nlocal), inner)
            xs.append((x, st["proj_vectors"], pj.xbar.copy() if pj.xbar is not None else None))
        outs[kind] = xs
    for (xa, na, ba), (xb, nb, bb) in zip(outs[1], outs[2]):
        assert na == nb
        assert np.max(np.abs(xa - xb)) <= 1e-10 * np.max(np.abs(xb))
        if ba is not None:
            assert np.max(np.abs(ba - bb)) <= 1e-9 * np.max(np.abs(bb))


def test_schedule_start_and_restart(orc):
    """From the start-th solve on the basis grows by one per solve and
    restarts at one vector after reaching max_vecs."""
    import oracle.projection as op
    mo, eo = _setup(orc)
    inner = _inner(orc, eo, tol=1e-8)
    pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=2, max_vecs=3, start=3)
    used = []
    for k in range(9):
        _, st = op.solve(eo, pj, _rhs(mo, 0.1 * k), np.zeros(mo.Nlocal), inner)
        used.append(st["proj_vectors"])
    assert used == [0, 0, 0, 1, 2, 3, 1, 2, 3]


def test_dependent_vector_is_discarded(orc):
    """A zero correction has no A-norm: the test (0/0) fails, the vector is
    dropped and the basis resets to x = xbar (:222-234); a zero xbar is dropped
    too and the next solve starts a new basis."""
    import oracle.projection as op
    mo, eo = _setup(orc)
    inner = _inner(orc, eo, tol=1e-10)
    pj = op.SolutionProjection(_A(eo), eo.invDegree, kind=2, max_vecs=8, start=1)
    for k in range(2):
        op.solve(eo, pj, _rhs(mo, 0.4 * k), np.zeros(mo.Nlocal), inner)
    assert len(pj.xx) == 2
    rhs = _rhs(mo, 0.4)                           # already in the span: xbar = x
    r = op.residual(eo, rhs, np.zeros(mo.Nlocal))
    pj.pre(r)
    xbar = pj.xbar.copy()
    x = pj.post(np.zeros(mo.Nlocal))              # zero correction
    assert pj.dropped == 1 and len(pj.xx) == 1
    assert np.max(np.abs(x - xbar)) == 0.0
    nA = pj.dot(pj.xx[0], eo.operator(pj.xx[0]))
    assert abs(nA - 1.0) <= 1e-12
    pj.pre(np.zeros(mo.Nlocal))                   # zero residual: xbar = 0
    pj.post(np.zeros(mo.Nlocal))
    assert pj.dropped == 3 and len(pj.xx) == 0
    _, st = op.solve(eo, pj, rhs, np.zeros(mo.Nlocal), inner)
    assert st["proj_vectors"] == 0 and len(pj.xx) == 1
