"""GPU parity of the projection initial guess (INITIAL GUESS = PROJECTION /
PROJECTION-ACONJ, SolutionProjection) against oracle.projection over a
sequence of solves with slowly varying right-hand sides, as in time
stepping.  Bars per solve: the basis size used matches exactly, the residual
before the projection to 1e-12 relative and after it to 1e-6 relative plus
1e-8 of the unprojected one, iterations +-1, solution 1e-8 relative."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _rhs(mo, t, nfields=1):
    out = []
    for f in range(nfields):
        out.append(np.sin(np.pi * mo.x + t + f) * np.cos(np.pi * mo.y - 0.5 * t) * (1 + mo.z * t)
                   + 0.01 * t * t * mo.x * mo.y)
    return np.concatenate(out)


def _sequence(gpu, torch, orc, kind, solver, ts, nfields=1, max_vecs=4, start=2, lam1=1.5):
    import oracle.mg as omg
    import oracle.projection as op
    shape, N = (4, 3, 3), 4
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    guess = {1: "projection", 2: "projection-aconj"}[kind]
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=lam1, poisson=False, nfields=nfields,
                      solver=solver, precon="jacobi", tol=1e-9, max_iter=2000, gmresRestart=12,
                      initial_guess=guess, proj_vectors=max_vecs, proj_start=start)
    eo = orc.Elliptic(mo, 1.0, lam1, False)
    inv = eo.inv_diag()
    L = mo.Nlocal
    if solver == "pcg":
        inner = lambda r: omg.pcg(eo, r, lambda v: inv * v, False, 1e-9, 2000)
    else:
        inner = lambda r: omg.gmres(eo, r, lambda v: inv * v, False, 12, 1e-9, 2000, nfields)
    Aop = lambda v: np.concatenate([eo.operator(v[f * L:(f + 1) * L]) for f in range(nfields)])
    pj = op.SolutionProjection(Aop, np.tile(eo.invDegree, nfields), kind=kind, max_vecs=max_vecs,
                               start=start)
    for k, t in enumerate(ts):
        rhs = _rhs(mo, t, nfields) if t is not None else np.zeros(nfields * L)
        xo, so = op.solve(eo, pj, rhs, np.zeros(nfields * L), inner, nfields)
        x = torch.zeros(nfields * L, dtype=torch.float64, device="cuda")
        sg = eg.solve(torch.from_numpy(rhs).cuda(), x)
        print("solve", k, "basis", sg["proj_vectors"], so["proj_vectors"], "it", sg["iterations"],
              so["iterations"], "res00", sg["res00"], "res0", sg["res0"], so["res0"], flush=True)
        assert sg["proj_vectors"] == so["proj_vectors"]
        assert abs(sg["res00"] - so["res00"]) <= 1e-12 * max(so["res00"], 1e-300)
        assert abs(sg["res0"] - so["res0"]) <= 1e-8 * max(so["res00"], 1e-300) + 1e-6 * so["res0"]
        assert abs(sg["iterations"] - so["iterations"]) <= 1
        if np.max(np.abs(xo)) > 0:
            assert rel(x.cpu().numpy(), xo) <= 1e-8
        else:
            assert float(x.abs().max()) == 0.0
    return pj


@pytest.mark.parametrize("kind", [1, 2])
def test_projection_pcg_sequence(gpu, torch_mod, orc, kind):
    """Ten solves with start 2, four vectors: warm-up, growth to the full basis,
    restarts; the projected residual drops well below the unprojected one."""
    ts = [0.05 * k for k in range(10)]
    _sequence(gpu, torch_mod, orc, kind, "pcg", ts)


def test_projection_reduces_iterations(gpu, torch_mod, orc):
    """With an absolute tolerance (as the pressure solve is usually run) the
    projected guess of a slowly varying right-hand side saves iterations."""
    torch = torch_mod
    mg = gpu.BoxMesh(4, 3, 3, 4, "kershaw", 0.3)
    its = {}
    for guess in ("previous", "projection-aconj"):
        eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=1.5, poisson=False, precon="jacobi", tol=1e-8,
                          relative=False, max_iter=2000, initial_guess=guess, proj_vectors=8,
                          proj_start=1)
        its[guess] = []
        for k in range(6):
            rhs = _rhs(mg, 0.02 * k)
            x = torch.zeros(rhs.size, dtype=torch.float64, device="cuda")
            st = eg.solve(torch.from_numpy(rhs).cuda(), x)
            assert st["converged"]
            its[guess].append(st["iterations"])
    print(its)
    assert its["projection-aconj"][0] == its["previous"][0]
    assert sum(its["projection-aconj"][2:]) < 0.7 * sum(its["previous"][2:])


def test_projection_gmres_discard_path(gpu, torch_mod, orc):
    """PGMRES with a zero right-hand side in the sequence: the zero correction
    fails the A-norm test, the basis resets, and a zero xbar is dropped too."""
    ts = [0.0, 0.1, 0.2, None, 0.3, 0.4]
    pj = _sequence(gpu, torch_mod, orc, 2, "gmres", ts, start=1, max_vecs=8)
    assert pj.dropped >= 2


def test_projection_block_three_fields(gpu, torch_mod, orc):
    ts = [0.1 * k for k in range(5)]
    _sequence(gpu, torch_mod, orc, 1, "gmres", ts, nfields=3, max_vecs=3, start=1, lam1=4.0)


def test_projection_rejects_bad_vector_count(gpu):
    mg = gpu.BoxMesh(2, 2, 2, 3)
    with pytest.raises(gpu.NrsError):
        gpu.Elliptic(mg, precon="jacobi", initial_guess="projection", proj_vectors=33)
