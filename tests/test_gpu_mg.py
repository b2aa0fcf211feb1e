"""GPU parity of the p-multigrid preconditioner (levels, Chebyshev-Jacobi
smoothing, transfers, FP64 coarse PCG) and of MG-preconditioned flexible CG
against the oracle (oracle/mg.py).  Bars: lambda_max(S A) per level within
1e-5 relative (FP32 operator inside the Arnoldi process), one V-cycle within
2e-4 relative in FP32 and 1e-9 in FP64, solves within +-1 iteration and 1e-8
relative in the solution; graph replays bit-identical to eager launches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def setup_pair(gpu, orc, n, N, levels, mg_precision=0, **kw):
    import oracle.mg as omg
    mg_mesh = gpu.BoxMesh(n, n, n, N, "kershaw", 0.3)
    eg = gpu.Elliptic(mg_mesh, precon="multigrid", solver="fcg", levels=levels,
                      mgPrecision=mg_precision, **kw)
    mo = orc.BoxMesh(n, n, n, N, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    T = np.float32 if mg_precision == 0 else np.float64
    om = omg.MG(mo, levels, T=T, degree=kw.get("cheby_degree", 2),
                coarse_tol=kw.get("coarse_tol", 1e-2), coarse_maxit=kw.get("coarse_max_iter", 200))
    return eg, eo, om


def assembled_random(eo, seed):
    r = np.random.default_rng(seed).uniform(-1, 1, eo.mesh.Nlocal)
    r[eo.maskIds] = 0
    return eo.ogs.gs(r, "add")


@pytest.mark.parametrize("prec", [0, 1])
def test_lambda_max_and_vcycle(gpu, torch_mod, orc, prec):
    torch = torch_mod
    eg, eo, om = setup_pair(gpu, orc, 4, 7, [7, 3, 1], mg_precision=prec)
    lv, _ = eg.mg_info()
    assert [l["order"] for l in lv] == [7, 3, 1]
    for l, L in zip(lv[:-1], om.levels[:-1]):
        assert abs(l["rho"] - L.rho) <= 1e-5 * L.rho, (l["rho"], L.rho)
    r = assembled_random(eo, 3)
    z = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    eg.preconditioner(dev(torch, r), z)
    torch.cuda.synchronize()
    zo = om.apply(r)
    tol = 2e-4 if prec == 0 else 1e-9
    assert rel(z.cpu().numpy(), zo) < tol


@pytest.mark.parametrize("levels", [[7, 3, 1], [7, 5, 3, 1], [5, 3, 1], [3, 1], [9, 5, 3, 1],
                                    [9, 3, 1]])
def test_mg_fcg_solve_vs_oracle(gpu, torch_mod, orc, levels):
    import oracle.mg as omg
    torch = torch_mod
    N = levels[0]
    eg, eo, om = setup_pair(gpu, orc, 4, N, levels, tol=1e-10, max_iter=400)
    rhs = np.random.default_rng(7).uniform(-1, 1, eg.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=400)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(levels, "gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8


def test_mg_graph_replay_matches_eager(gpu, torch_mod, orc):
    """The coarse PCG chunks replay a captured graph many times per solve with
    eager work in between; the solve must equal the all-eager one bit for bit."""
    torch = torch_mod
    m = gpu.BoxMesh(5, 5, 5, 7, "kershaw", 0.3)
    rhs = np.random.default_rng(1).uniform(-1, 1, m.Nlocal)
    out = []
    for graphs in (1, 0, 1):
        el = gpu.Elliptic(m, precon="multigrid", solver="fcg", levels=[7, 3, 1], tol=1e-9,
                          useGraphs=graphs)
        x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
        st = el.solve(dev(torch, rhs), x)
        st2_x = torch.zeros_like(x)
        el.solve(dev(torch, rhs), st2_x)              # second solve replays the same graph again
        out.append((st["iterations"], x.cpu().numpy(), st2_x.cpu().numpy()))
    assert out[0][0] == out[1][0] == out[2][0]
    for o in out[1:]:
        assert np.array_equal(o[1], out[0][1]) and np.array_equal(o[2], out[0][1])


def test_fp64_and_fp32_mg_reach_same_solution(gpu, torch_mod, orc):
    torch = torch_mod
    m = gpu.BoxMesh(6, 6, 6, 7, "kershaw", 0.3)
    rhs = np.random.default_rng(2).uniform(-1, 1, m.Nlocal)
    xs = []
    for prec in (0, 1):
        el = gpu.Elliptic(m, precon="multigrid", solver="fcg", levels=[7, 3, 1], tol=1e-10,
                          mgPrecision=prec)
        x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
        st = el.solve(dev(torch, rhs), x)
        assert st["converged"]
        xs.append(x.cpu().numpy())
    assert rel(xs[0], xs[1]) < 1e-8


def test_setup_on_recycled_memory(gpu, torch_mod):
    """Work vectors freed by a solve that went NaN are reused by the next
    setup's allocations; nothing may read them before writing (regression: the
    Arnoldi start used y = a x + 0 y on a fresh allocation)."""
    torch = torch_mod
    m = gpu.BoxMesh(4, 4, 4, 5, "kershaw", 0.3)
    bad = np.full(m.Nlocal, np.nan)
    for _ in range(2):
        el = gpu.Elliptic(m, precon="multigrid", solver="fcg", levels=[5, 3, 1])
        with pytest.raises(gpu.NrsError):
            el.solve(dev(torch, bad), torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda"))
        del el
    el = gpu.Elliptic(m, precon="multigrid", solver="fcg", levels=[5, 3, 1], tol=1e-10)
    lv, _ = el.mg_info()
    assert all(np.isfinite(l["rho"]) for l in lv)
    rhs = np.random.default_rng(7).uniform(-1, 1, m.Nlocal)
    st = el.solve(dev(torch, rhs), torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda"))
    assert st["converged"]


@pytest.mark.parametrize("shape,N,smoother,lam", [((10, 9, 8), 3, "chebyshev+jacobi", 0.0),
                                                  ((9, 8, 8), 5, "opt4chebyshev+ras", 2.0)])
def test_mg_amg_coarse_vs_oracle(gpu, torch_mod, orc, shape, N, smoother, lam):
    """Coarse level by FP64 PCG on the assembled order-1 matrix with one
    smoothed-aggregation V-cycle as preconditioner (more than 256 coarse
    unknowns, so the hierarchy has several levels): outer iterations within
    +-1 of the oracle, coarse iterations within +-1, solution 1e-8."""
    import oracle.mg as omg
    torch = torch_mod
    eps = 0.3
    poisson = lam == 0.0
    mg = gpu.BoxMesh(*shape, N, "kershaw", eps)
    S = "ras" if "ras" in smoother else "jacobi"
    cheby = "opt4" if smoother.startswith("opt4") else "cheby1"
    deg = 3 if cheby == "opt4" else 2
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=lam, poisson=poisson, precon="multigrid",
                      solver="fcg", levels=[N, 1], smoother=smoother, cheby_degree=deg,
                      coarse="amg", tol=1e-10, max_iter=200)
    mo = orc.BoxMesh(*shape, N, "kershaw", eps)
    eo = orc.Elliptic(mo, 1.0, lam, poisson)
    om = omg.MG(mo, [N, 1], T=np.float32, degree=deg, cheby=cheby, S=S, coarse="amg",
                lambda0=1.0, lambda1=lam, poisson=poisson)
    assert len(om.levels[-1].amg.levels) >= 2
    rhs = np.random.default_rng(N).uniform(-1, 1, mo.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros(mo.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(torch.from_numpy(rhs).cuda(), x)
    _, cit = eg.mg_info()
    print(shape, N, smoother, "gpu it", st["iterations"], "oracle it", info["iterations"],
          "coarse it", cit, om.last_coarse_it)
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert abs(cit - om.last_coarse_it) <= 1
    xh = x.cpu().numpy()
    assert np.max(np.abs(xh - xo)) <= 1e-8 * np.max(np.abs(xo))


def test_mg_amg_fixed_iterations_graph_vs_oracle(gpu, torch_mod, orc):
    """AMG coarse solve with a fixed iteration count (coarseTol = 0, two
    AMG-preconditioned PCG iterations, no host round trip): the whole V-cycle
    is replayed as one HIP graph; the result equals the eager cycle bit for
    bit and the solve matches the oracle (iterations +-1, solution 1e-8)."""
    import oracle.mg as omg
    torch = torch_mod
    shape, N = (10, 9, 8), 3
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    kw = dict(precon="multigrid", solver="fcg", levels=[N, 1], smoother="opt4chebyshev+ras",
              cheby_degree=3, fdmPrecision=1, coarse="amg", coarse_tol=0.0, coarse_max_iter=2,
              tol=1e-10, max_iter=200)
    eg = gpu.Elliptic(mg, **kw)
    ee = gpu.Elliptic(mg, useGraphs=0, **kw)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    om = omg.MG(mo, [N, 1], T=np.float32, degree=3, cheby="opt4", S="ras", fmt="bf16",
                coarse="amg", coarse_tol=0.0, coarse_maxit=2)
    rhs = np.random.default_rng(4).uniform(-1, 1, mo.Nlocal)
    r = torch.from_numpy(rhs).cuda()
    z1 = torch.zeros_like(r)
    z2 = torch.zeros_like(r)
    for _ in range(3):                      # replays after the capture
        eg.preconditioner(r, z1)
        ee.preconditioner(r, z2)
        torch.cuda.synchronize()
        assert torch.equal(z1, z2)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros_like(r)
    st = eg.solve(r.clone(), x)
    print("fixed AMG gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"] and abs(st["iterations"] - info["iterations"]) <= 1
    assert np.max(np.abs(x.cpu().numpy() - xo)) <= 1e-8 * np.max(np.abs(xo))


def test_mg_schedule_degrees_vs_oracle(gpu, torch_mod, orc):
    """A MULTIGRID SCHEDULE with a different Chebyshev degree per level and leg
    (parsed by the library, levels and degrees as the reference's level setup
    reads them) against the oracle's V-cycle with the same degrees: one V-cycle
    within 2e-4 (FP32), FCG within +-1 iteration and 1e-8; and a schedule with
    degree 2 everywhere gives the plain levels=[7,3,1], degree-2 V-cycle bit for bit."""
    import oracle.mg as omg
    torch = torch_mod
    sched = "p=7+degree=2, p=3+degree=1, p=1, p=3+degree=4, p=7+degree=3"
    lv, pre, post, errs = omg.parse_mg_schedule(sched, 7, 2)
    assert not errs and lv == [7, 3, 1] and pre == [3, 4, 0] and post == [2, 1, 0]
    mesh = gpu.BoxMesh(4, 4, 4, 7, "kershaw", 0.3)
    eg = gpu.Elliptic(mesh, precon="multigrid", solver="fcg", schedule=sched, cheby_degree=2,
                      tol=1e-10, max_iter=400)
    assert [l["order"] for l in eg.mg_info()[0]] == [7, 3, 1]
    mo = orc.BoxMesh(4, 4, 4, 7, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    om = omg.MG(mo, lv, T=np.float32, degree=2, pre_degree=pre, post_degree=post)
    r = assembled_random(eo, 5)
    z = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    eg.preconditioner(dev(torch, r), z)
    torch.cuda.synchronize()
    assert rel(z.cpu().numpy(), om.apply(r)) < 2e-4
    rhs = np.random.default_rng(8).uniform(-1, 1, eg.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=400)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print("schedule gpu it", st["iterations"], "oracle it", info["iterations"], flush=True)
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8
    # uniform schedule == plain options
    e1 = gpu.Elliptic(mesh, precon="multigrid", solver="fcg", cheby_degree=2, autotune=0,
                      schedule="p=7+degree=2, p=3+degree=2, p=1, p=3+degree=2, p=7+degree=2")
    e2 = gpu.Elliptic(mesh, precon="multigrid", solver="fcg", cheby_degree=2, autotune=0,
                      levels=[7, 3, 1])
    z1 = torch.zeros_like(z)
    z2 = torch.zeros_like(z)
    e1.preconditioner(dev(torch, r), z1)
    e2.preconditioner(dev(torch, r), z2)
    torch.cuda.synchronize()
    assert torch.equal(z1, z2)


@pytest.mark.parametrize("mgp,smoother", [(0, "opt4chebyshev+ras"), (1, "chebyshev+jacobi")])
def test_vcycle_from_level_operations(gpu, torch_mod, orc, mgp, smoother):
    """A V-cycle composed by hand from the per-level operations
    (nrs_elliptic_mg_level: smooth, residual, coarsen, prolongate, as
    MGSolver_t::runVcycle, MGSolver.cpp:167-193, calls them) gives the library's
    preconditioner bit for bit; the level Ax and residual agree with each other."""
    torch = torch_mod
    mesh = gpu.BoxMesh(3, 3, 2, 5, "kershaw", 0.3)
    el = gpu.Elliptic(mesh, precon="multigrid", solver="fcg", levels=[5, 3, 1], smoother=smoother,
                      cheby_degree=2, coarse="smooth_jacobi", mgPrecision=mgp, fdmPrecision=0,
                      autotune=0)
    T = torch.float32 if mgp == 0 else torch.float64
    lv, _ = el.mg_info()
    n = [l["Nlocal"] for l in lv]
    eo = orc.Elliptic(orc.BoxMesh(3, 3, 2, 5, "kershaw", 0.3))
    r = dev(torch, assembled_random(eo, 9))
    z = torch.zeros_like(r)
    el.preconditioner(r, z)
    rhs = [r.to(T)] + [torch.zeros(n[k], dtype=T, device="cuda") for k in (1, 2)]
    x = [torch.zeros(n[k], dtype=T, device="cuda") for k in range(3)]
    res = [torch.zeros(n[k], dtype=T, device="cuda") for k in range(3)]
    for k in (0, 1):
        el.mg_level(k, "smooth", rhs[k], x[k], x_is_zero=True)
        el.mg_level(k, "residual", rhs[k], res[k], x=x[k])
        el.mg_level(k + 1, "coarsen", res[k], rhs[k + 1])
    el.mg_level(2, "smooth", rhs[2], x[2], x_is_zero=True)
    for k in (1, 0):
        el.mg_level(k + 1, "prolongate", x[k + 1], x[k])
        el.mg_level(k, "smooth", rhs[k], x[k], x_is_zero=False)
    torch.cuda.synchronize()
    assert torch.equal(x[0].to(torch.float64), z)
    # residual = rhs - A x through the level Ax
    ax = torch.zeros_like(x[1])
    el.mg_level(1, "ax", x[1], ax)
    rr = torch.zeros_like(x[1])
    el.mg_level(1, "residual", rhs[1], rr, x=x[1])
    torch.cuda.synchronize()
    assert torch.equal(rr, rhs[1] - ax)
    with pytest.raises(gpu.NrsError):
        el.mg_level(0, "coarsen", res[0], rhs[1])          # level 0 has no finer level
