"""GPU parity of the solver-level entry points (ellipticSolveSetup state,
ellipticOperator, ellipticSolve with PCG / flexible CG) against the oracle.
Bars: FP64 operator <= 1e-12 relative; solves agree with the oracle in
iteration count (+-1) and solution (<= 1e-8 relative), and reach the requested
true relative residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


def both(gpu, orc, nx, ny, nz, N, mp, eps=0.3, **kw):
    mg = gpu.BoxMesh(nx, ny, nz, N, mp, eps)
    mo = orc.BoxMesh(nx, ny, nz, N, mp, eps)
    return gpu.Elliptic(mg, **kw), orc.Elliptic(mo)


@pytest.mark.parametrize("N", [1, 3, 5, 7, 9])
def test_setup_state_matches_oracle(gpu, torch_mod, orc, N):
    eg, eo = both(gpu, orc, 3, 2, 2, N, "kershaw")
    assert eg.Nmasked == len(eo.maskIds)
    assert np.array_equal(host(eg.mask_ids), eo.maskIds.astype(np.int32))
    assert np.array_equal(host(eg.inv_degree), eo.invDegree)
    assert abs(eg.volume - eo.volume) < 1e-13
    ggeo = host(eg.ggeo).reshape(-1, 7, (N + 1) ** 3)
    ref = eo.mesh.ggeo.reshape(-1, 7, (N + 1) ** 3)
    assert np.max(np.abs(ggeo - ref)) < 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("N", [2, 3, 7])
@pytest.mark.parametrize("mp", ["kershaw", "sine"])
def test_operator_vs_oracle(gpu, torch_mod, orc, N, mp):
    torch = torch_mod
    eg, eo = both(gpu, orc, 4, 3, 2, N, mp)
    q = np.random.default_rng(N).uniform(-1, 1, eg.Nlocal)
    ref = eo.operator(q)
    Aq = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    eg.operator(dev(torch, q), Aq)
    torch.cuda.synchronize()
    assert rel(host(Aq), ref) < 1e-12
    # masked nodes are zero, assembled values are continuous
    out = host(Aq)
    assert np.all(out[eo.maskIds] == 0)
    assert np.array_equal(eo.ogs.gs(out, "max"), out)


def _manufactured(mesh, ggeo):
    Np = (mesh.N + 1) ** 3
    gw = ggeo.reshape(-1, 7, Np)[:, 6].ravel()
    u = np.sin(np.pi * mesh.x) * np.sin(np.pi * mesh.y) * np.sin(np.pi * mesh.z)
    return u, gw * 3 * np.pi ** 2 * u


@pytest.mark.parametrize("solver", ["pcg", "fcg"])
@pytest.mark.parametrize("precon", ["jacobi", "none"])
def test_solve_small_vs_oracle(gpu, torch_mod, orc, solver, precon):
    torch = torch_mod
    eg, eo = both(gpu, orc, 3, 3, 3, 5, "sine", precon=precon, solver=solver, tol=1e-8, max_iter=800)
    u, rhs = _manufactured(eo.mesh, eo.mesh.ggeo)
    xo, info = eo.solve(rhs, precon=precon, flexible=(solver == "fcg"), tol=1e-8, maxit=800)
    r = dev(torch, rhs)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(r, x)
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert abs(st["res0"] - info["res0"]) <= 1e-12 * info["res0"]
    if precon == "jacobi":
        assert st["converged"]
        assert rel(host(x), xo) <= 1e-8


def test_solve_nonzero_initial_guess(gpu, torch_mod, orc):
    torch = torch_mod
    eg, eo = both(gpu, orc, 3, 2, 2, 4, "kershaw", tol=1e-10, max_iter=500)
    rng = np.random.default_rng(9)
    rhs = rng.uniform(-1, 1, eg.Nlocal)
    x0 = eo.ogs.gs(rng.uniform(-1, 1, eg.Nlocal), "max")   # continuous
    x0[eo.maskIds] = 0.0
    xo, info = eo.solve(rhs, x0=x0, tol=1e-10, maxit=500)
    x = dev(torch, x0)
    st = eg.solve(dev(torch, rhs), x)
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(host(x), xo) <= 1e-8


def test_c0_plain_poisson(gpu, torch_mod, orc):
    """8x8x8 elements, N=7, sine-deformed, Jacobi PCG to 1e-8 relative, FP64."""
    torch = torch_mod
    eg, eo = both(gpu, orc, 8, 8, 8, 7, "sine", tol=1e-8, max_iter=1000)
    u, rhs = _manufactured(eo.mesh, eo.mesh.ggeo)
    xo, info = eo.solve(rhs, tol=1e-8, maxit=1000)
    r = dev(torch, rhs)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(r, x)
    xg = host(x)
    print("C0: iterations gpu %d oracle %d, res0 %.6e res %.6e, %.3f s" %
          (st["iterations"], info["iterations"], st["res0"], st["res"], st["seconds"]))
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(xg, xo) <= 1e-8
    # true residual of the returned solution, in the solver's norm
    Ax = torch.zeros_like(x)
    eg.operator(x, Ax)
    b = eo.ogs.gs(rhs, "add")
    b[eo.maskIds] = 0
    res_true = np.sqrt(np.sum(eo.invDegree * (b - host(Ax)) ** 2) / eo.volume)
    assert res_true <= 1.05e-8 * st["res0"]
    # discretisation error against the manufactured solution
    assert np.max(np.abs(xg - u)) < 1e-5


def test_solve_nan_rhs_reports_error(gpu, torch_mod, orc):
    torch = torch_mod
    mg = gpu.BoxMesh(2, 2, 2, 3)
    eg = gpu.Elliptic(mg)
    eo = orc.Elliptic(orc.BoxMesh(2, 2, 2, 3))
    rhs = np.ones(eg.Nlocal)
    rhs[1 + 4 * (1 + 4 * 1)] = np.nan          # node (1,1,1) of element 0: not on a Dirichlet face
    _, info = eo.solve(rhs)
    assert info["iterations"] < 0 or np.isnan(info["res0"])   # the oracle sees the NaN as well
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.NrsError) as ex:
        eg.solve(dev(torch, rhs), x)
    assert ex.value.code == -3
    # a NaN on a Dirichlet node is masked away before the first norm (mask before gs)
    rhs = np.ones(eg.Nlocal)
    rhs[5] = np.nan
    st = eg.solve(dev(torch, rhs), torch.zeros_like(x))
    assert st["converged"]


def test_solve_max_iter_not_converged(gpu, torch_mod, orc):
    torch = torch_mod
    eg, eo = both(gpu, orc, 3, 3, 3, 5, "sine", tol=1e-12, max_iter=3)
    u, rhs = _manufactured(eo.mesh, eo.mesh.ggeo)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    assert st["iterations"] == 3 and not st["converged"]


def test_c1_shape_properties(gpu, torch_mod):
    """32^3 elements, N=7, Kershaw 0.3: size-independent properties of the
    FP64 operator at the benchmark size (symmetry, constants in the kernel
    of the unmasked element apply, linearity)."""
    torch = torch_mod
    m = gpu.BoxMesh(32, 32, 32, 7, "kershaw", 0.3)
    el = gpu.Elliptic(m, precon="none")
    L = el.Nlocal
    g = torch.Generator(device="cuda").manual_seed(0)
    ogs = gpu.Ogs(m.gid)
    p = torch.rand(L, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    q = torch.rand(L, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    ogs.gather_scatter(p, "max"); ogs.gather_scatter(q, "max")     # continuous fields
    w = el.inv_degree
    mask = el.mask_ids.long()
    p[mask] = 0; q[mask] = 0                                       # in the masked space
    Ap, Aq = torch.empty_like(p), torch.empty_like(q)
    el.operator(p, Ap); el.operator(q, Aq)
    a, b = torch.sum(w * q * Ap).item(), torch.sum(w * p * Aq).item()
    assert abs(a - b) <= 1e-11 * abs(a)
    Apq = torch.empty_like(p)
    el.operator(2.0 * p - 3.0 * q, Apq)
    assert (torch.max(torch.abs(Apq - (2.0 * Ap - 3.0 * Aq))) / torch.max(torch.abs(Apq))).item() < 1e-12
    one = torch.ones(L, dtype=torch.float64, device="cuda")
    A1 = torch.empty_like(one)
    el.operator(one, A1, masked=False)
    # constants are in the kernel of every element apply, so the assembled result is 0
    assert (torch.max(torch.abs(A1)) / torch.max(torch.abs(Ap))).item() < 1e-12


def _block_pcg(eo, rhs3, tol, maxit, precon=None, flexible=False):
    """(Flexible) PCG over three fields with one operator (ellipticSolve /
    PCG.cpp:85-203 with Nfields = 3): inner products summed over the fields
    with invDegree weights, norm scaled by 1/volume, relative tolerance;
    Jacobi unless precon (one call per field) is given."""
    L = rhs3.size // 3
    w = np.tile(eo.invDegree, 3)
    inv = np.tile(eo.inv_diag(), 3)
    M = (lambda v: inv * v) if precon is None else \
        (lambda v: np.concatenate([precon(v[f * L:(f + 1) * L]) for f in range(3)]))
    A = lambda v: np.concatenate([eo.operator(v[f * L:(f + 1) * L]) for f in range(3)])
    r = np.concatenate([eo.ogs.gs(np.where(np.isin(np.arange(L), eo.maskIds), 0.0,
                                           rhs3[f * L:(f + 1) * L]), "add") for f in range(3)])
    x = np.zeros_like(r)
    res0 = np.sqrt(np.sum(w * r * r) / eo.volume)
    tol = tol * res0
    p = np.zeros_like(r)
    Ap = np.zeros_like(r)
    rdotz_old, alpha, it = 0.0, 0.0, 0
    while it < maxit:
        z = M(r)
        rdotz = np.sum(w * r * z)
        if it == 0:
            beta = 0.0
        elif flexible:             # PCG.cpp:141-147: beta = -alpha <z, Ap_old> / <r,z>_old
            beta = -alpha * np.sum(w * z * Ap) / rdotz_old
        else:
            beta = rdotz / rdotz_old
        p = z + beta * p
        Ap = A(p)
        alpha = rdotz / (np.sum(w * p * Ap) + 1e-300)
        r = r - alpha * Ap
        x = x + alpha * p
        rdotz_old = rdotz
        it += 1
        if np.sqrt(np.sum(w * r * r) / eo.volume) <= tol:
            break
    return x, it


@pytest.mark.parametrize("N,lam1", [(3, 10.0), (5, 0.0), (7, 2.5)])
def test_block_three_fields_vs_oracle(gpu, torch_mod, orc, N, lam1):
    """Block (velocity-type) solve of three fields sharing one Helmholtz
    operator: the block operator equals the oracle's operator on every field
    (1e-12); the block Jacobi-PCG matches the oracle's three-field PCG in
    iterations (+-1) and solution (1e-8)."""
    torch = torch_mod
    shape = (4, 3, 3)
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=lam1, poisson=False, nfields=3, precon="jacobi",
                      tol=1e-9, max_iter=1000)
    eo = orc.Elliptic(mo, 1.0, lam1, False)
    L = mo.Nlocal
    assert eg.Nlocal == L
    q = np.random.default_rng(N).uniform(-1, 1, 3 * L)
    Aq = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
    eg.operator(dev(torch, q), Aq)
    torch.cuda.synchronize()
    got = host(Aq)
    for f in range(3):
        ref = eo.operator(q[f * L:(f + 1) * L])
        assert rel(got[f * L:(f + 1) * L], ref) < 1e-12
    rhs = np.random.default_rng(N + 1).uniform(-1, 1, 3 * L)
    xo, it = _block_pcg(eo, rhs, 1e-9, 1000)
    x = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(N, lam1, "gpu it", st["iterations"], "oracle it", it)
    assert st["converged"] and abs(st["iterations"] - it) <= 1
    assert rel(host(x), xo) <= 1e-8


def test_block_rejects_two_fields(gpu):
    m = gpu.BoxMesh(2, 2, 2, 3)
    with pytest.raises(gpu.NrsError):
        gpu.Elliptic(m, nfields=2, precon="jacobi")


def test_block_multigrid_vs_oracle(gpu, torch_mod, orc):
    """Three fields with the p-multigrid hierarchy applied to each field and
    flexible CG over the block: iterations +-1 and solution 1e-8 against the
    oracle's block FCG with its multigrid per field."""
    import oracle.mg as omg
    torch = torch_mod
    shape, N, lam1 = (4, 3, 3), 5, 3.0
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=lam1, poisson=False, nfields=3, solver="fcg",
                      precon="multigrid", levels=[N, 3, 1], smoother="opt4chebyshev+jacobi",
                      cheby_degree=3, coarse="smooth", tol=1e-9, max_iter=200)
    eo = orc.Elliptic(mo, 1.0, lam1, False)
    om = omg.MG(mo, [N, 3, 1], T=np.float32, degree=3, cheby="opt4", S="jacobi", coarse="smooth",
                lambda0=1.0, lambda1=lam1, poisson=False)
    L = mo.Nlocal
    rhs = np.random.default_rng(9).uniform(-1, 1, 3 * L)
    xo, it = _block_pcg(eo, rhs, 1e-9, 200, precon=om.apply, flexible=True)
    x = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print("block MG gpu it", st["iterations"], "oracle it", it)
    assert st["converged"] and abs(st["iterations"] - it) <= 1
    assert rel(host(x), xo) <= 1e-8


@pytest.mark.parametrize("solver,precon,restart", [("fgmres", "multigrid", 15), ("gmres", "jacobi", 15),
                                                   ("gmres", "jacobi", 5), ("fgmres", "none", 8)])
def test_pgmres_vs_oracle(gpu, torch_mod, orc, solver, precon, restart):
    """Restarted (flexible) PGMRES (PGMRES.cpp:111-340) against the oracle's
    restatement with the same preconditioner: iterations +-1, solution 1e-8;
    restart 5 and 8 exercise the restart path."""
    import oracle.mg as omg
    torch = torch_mod
    shape, N, lam1 = (4, 3, 3), 5, 1.5
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    kw = dict(solver=solver, precon=precon, gmresRestart=restart, tol=1e-9, max_iter=600)
    if precon == "multigrid":
        kw.update(levels=[N, 3, 1], smoother="opt4chebyshev+jacobi", cheby_degree=3, coarse="smooth")
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=lam1, poisson=False, **kw)
    eo = orc.Elliptic(mo, 1.0, lam1, False)
    if precon == "multigrid":
        om = omg.MG(mo, [N, 3, 1], T=np.float32, degree=3, cheby="opt4", S="jacobi",
                    coarse="smooth", lambda0=1.0, lambda1=lam1, poisson=False)
        M = om.apply
    elif precon == "jacobi":
        inv = eo.inv_diag()
        M = lambda r: inv * r
    else:
        M = lambda r: r
    rhs = np.random.default_rng(restart).uniform(-1, 1, mo.Nlocal)
    xo, info = omg.gmres_solve(eo, rhs, M, flexible=(solver == "fgmres"), restart=restart,
                               tol=1e-9, maxit=600)
    x = torch.zeros(mo.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(solver, precon, restart, "gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"] and abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(host(x), xo) <= 1e-8


def test_pgmres_block_three_fields(gpu, torch_mod, orc):
    import oracle.mg as omg
    torch = torch_mod
    shape, N = (4, 3, 3), 3
    mg = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    eg = gpu.Elliptic(mg, lambda0=1.0, lambda1=4.0, poisson=False, nfields=3, solver="gmres",
                      precon="jacobi", gmresRestart=10, tol=1e-9, max_iter=600)
    eo = orc.Elliptic(mo, 1.0, 4.0, False)
    inv = eo.inv_diag()
    rhs = np.random.default_rng(3).uniform(-1, 1, 3 * mo.Nlocal)
    xo, info = omg.gmres_solve(eo, rhs, lambda r: inv * r, flexible=False, restart=10, tol=1e-9,
                               maxit=600, nfields=3)
    x = torch.zeros(3 * mo.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    assert st["converged"] and abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(host(x), xo) <= 1e-8
