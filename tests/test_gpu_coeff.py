"""Coefficient fields (ELLIPTIC COEFF FIELD = TRUE: per-node lambda0 / lambda1,
elliptic_t::o_lambda0 / o_lambda1) and the per-solve Jacobi refresh (ELLIPTIC
PRECO COEFF FIELD, ellipticSolve.cpp:79-88 -> ellipticUpdateJacobi) against the
oracle's field operator (oracle.FieldElliptic: ax and build_diag with fields,
both pinned against the reference's compiled kernels and by brute force).

Bars: operator 1e-12 relative; solves iterations +-1 and solution 1e-8
relative.  The velocity case is the reference's default velocity solve: block
(three-field) Jacobi-PCG with lambda0 = mu(x) and lambda1 = rho(x) bd0 / dt."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def fields(m, seed, mu0=0.01, rho_bd0_dt=1500.0):
    """A smooth viscosity mu(x) > 0 and rho(x) bd0 / dt > 0 at the nodes."""
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.2, 0.5, 2)
    mu = mu0 * (1.0 + a * np.sin(2 * np.pi * m.x) * np.cos(np.pi * m.y) + 0.3 * m.z)
    h = rho_bd0_dt * (1.0 + b * np.cos(np.pi * m.x * m.y) + 0.2 * np.sin(3 * m.z))
    return mu, h


def block_pcg(eo, rhs3, tol, maxit):
    """Jacobi-PCG over three fields with one operator (ellipticSolve / PCG.cpp:85-203,
    Nfields = 3): weighted inner products summed over the fields, relative tolerance."""
    L = rhs3.size // 3
    w = np.tile(eo.invDegree, 3)
    inv = np.tile(eo.inv_diag(), 3)
    A = lambda v: np.concatenate([eo.operator(v[f * L:(f + 1) * L]) for f in range(3)])
    mk = np.isin(np.arange(L), eo.maskIds)
    r = np.concatenate([eo.ogs.gs(np.where(mk, 0.0, rhs3[f * L:(f + 1) * L]), "add") for f in range(3)])
    x = np.zeros_like(r)
    res0 = np.sqrt(np.sum(w * r * r) / eo.volume)
    tol = tol * res0
    p = np.zeros_like(r)
    rdotz_old, it = 0.0, 0
    while it < maxit:
        z = inv * r
        rdotz = np.sum(w * r * z)
        beta = 0.0 if it == 0 else rdotz / rdotz_old
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


@pytest.mark.parametrize("N,poisson", [(7, True), (7, False), (5, False)])
def test_field_operator_vs_oracle(gpu, torch_mod, orc, N, poisson):
    """The operator with fields, on a handle set up with the default options
    (at N = 7 the constant-coefficient operator would take the face fusion)."""
    torch = torch_mod
    mg = gpu.BoxMesh(6, 3, 2, N, "kershaw", 0.3)
    mo = orc.BoxMesh(6, 3, 2, N, "kershaw", 0.3)
    mu, h = fields(mo, N)
    l0 = dev(torch, mu)
    l1 = None if poisson else dev(torch, h)
    eg = gpu.Elliptic(mg, poisson=poisson, lambda0_field=l0, lambda1_field=l1, precon="jacobi")
    eo = orc.FieldElliptic(mo, mu, None if poisson else h, poisson=poisson)
    q = np.random.default_rng(1).uniform(-1, 1, mo.Nlocal)
    Aq = torch.empty(mo.Nlocal, dtype=torch.float64, device="cuda")
    eg.operator(dev(torch, q), Aq)
    print("field operator N=%d poisson=%s rel err %.2e" % (N, poisson, rel(Aq.cpu().numpy(), eo.operator(q))),
          flush=True)
    assert rel(Aq.cpu().numpy(), eo.operator(q)) < 1e-12
    inv = eg.buffer(4)                      # invDiagA: the Jacobi diagonal with the fields
    assert rel(inv.cpu().numpy()[:mo.Nlocal], eo.inv_diag()) < 1e-12


def test_helmholtz_field_solve_vs_oracle(gpu, torch_mod, orc):
    import oracle.mg as omg
    torch = torch_mod
    N = 5
    mg = gpu.BoxMesh(4, 3, 3, N, "kershaw", 0.3)
    mo = orc.BoxMesh(4, 3, 3, N, "kershaw", 0.3)
    mu, h = fields(mo, 3, mu0=1.0, rho_bd0_dt=20.0)
    eg = gpu.Elliptic(mg, poisson=False, lambda0_field=dev(torch, mu), lambda1_field=dev(torch, h),
                      precon="jacobi", tol=1e-9, max_iter=1000)
    eo = orc.FieldElliptic(mo, mu, h, poisson=False)
    rhs = np.random.default_rng(4).uniform(-1, 1, mo.Nlocal)
    inv = eo.inv_diag()
    xo, info = omg.fcg_solve(eo, rhs, lambda r: inv * r, flexible=False, tol=1e-9, maxit=1000)
    x = torch.zeros(mo.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print("field Helmholtz: gpu %d iterations, oracle %d" % (st["iterations"], info["iterations"]), flush=True)
    assert st["converged"] and abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8


def test_velocity_block_solve_with_fields(gpu, torch_mod, orc):
    """The reference's default velocity solve (block Jacobi-PCG, three fields,
    one operator) with lambda0 = mu(x) and lambda1 = rho(x) bd0 / dt; then the
    fields change in place and the next solve, whose Jacobi diagonal is rebuilt
    from the new values, matches the oracle with the new fields."""
    torch = torch_mod
    N = 7
    mg = gpu.BoxMesh(4, 3, 3, N, "kershaw", 0.3)
    mo = orc.BoxMesh(4, 3, 3, N, "kershaw", 0.3)
    mu, h = fields(mo, 5)
    l0, l1 = dev(torch, mu), dev(torch, h)
    eg = gpu.Elliptic(mg, poisson=False, nfields=3, lambda0_field=l0, lambda1_field=l1, precon="jacobi",
                      tol=1e-10, max_iter=1000)
    L = mo.Nlocal
    rhs = np.random.default_rng(6).uniform(-1, 1, 3 * L)
    for step in range(2):
        if step == 1:                       # a new time step: new properties, same buffers
            mu, h = fields(mo, 11, mu0=0.02, rho_bd0_dt=900.0)
            l0.copy_(dev(torch, mu))
            l1.copy_(dev(torch, h))
        eo = orc.FieldElliptic(mo, mu, h, poisson=False)
        xo, it = block_pcg(eo, rhs, 1e-10, 1000)
        x = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
        st = eg.solve(dev(torch, rhs), x)
        print("velocity block, step %d: gpu %d iterations, oracle %d" % (step, st["iterations"], it), flush=True)
        assert st["converged"] and abs(st["iterations"] - it) <= 1
        assert rel(x.cpu().numpy(), xo) <= 1e-8


def test_refresh_switch_and_constants(gpu, torch_mod):
    """precoCoeffField = 0 keeps the Jacobi diagonal of the old values until
    update_jacobi is called, after which the solve equals the refreshing
    handle's bit for bit; set_coefficients(None, None) gives the constant
    operator bit for bit."""
    torch = torch_mod
    m = gpu.BoxMesh(4, 3, 2, 5, "kershaw", 0.3)
    mu, h = fields(m, 2, mu0=1.0, rho_bd0_dt=10.0)
    a0, a1 = dev(torch, mu), dev(torch, h)
    b0, b1 = dev(torch, mu), dev(torch, h)
    kw = dict(poisson=False, precon="jacobi", tol=1e-9, max_iter=500)
    fresh = gpu.Elliptic(m, lambda0_field=a0, lambda1_field=a1, **kw)
    stale = gpu.Elliptic(m, lambda0_field=b0, lambda1_field=b1, precoCoeffField=0, **kw)
    mu2, h2 = fields(m, 9, mu0=3.0, rho_bd0_dt=0.5)
    for t, v in ((a0, mu2), (a1, h2), (b0, mu2), (b1, h2)):
        t.copy_(dev(torch, v))
    rhs = dev(torch, np.random.default_rng(8).uniform(-1, 1, m.Nlocal))
    xs = []
    for el in (fresh, stale):
        x = torch.zeros_like(rhs)
        xs.append((el.solve(rhs.clone(), x), x))
    assert xs[0][0]["converged"]
    assert not torch.equal(xs[0][1], xs[1][1])        # stale diagonal: a different iteration
    stale.update_jacobi()
    x = torch.zeros_like(rhs)
    st = stale.solve(rhs.clone(), x)
    assert st["iterations"] == xs[0][0]["iterations"] and torch.equal(x, xs[0][1])
    # back to the descriptor's constants (lambda0 = 1, lambda1 = 0 for `fresh`)
    fresh.set_coefficients(None, None)
    const = gpu.Elliptic(m, lambda0=1.0, lambda1=0.0, **kw)
    q = dev(torch, np.random.default_rng(9).uniform(-1, 1, m.Nlocal))
    y0, y1 = torch.empty_like(q), torch.empty_like(q)
    fresh.operator(q, y0)
    const.operator(q, y1)
    assert torch.equal(y0, y1)
    assert torch.equal(fresh.buffer(4), const.buffer(4))


def test_multigrid_with_fields_refused(gpu, torch_mod):
    torch = torch_mod
    m = gpu.BoxMesh(2, 2, 2, 3)
    l0 = torch.ones(m.Nlocal, dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.NrsError):
        gpu.Elliptic(m, lambda0_field=l0, precon="multigrid", levels=[3, 1])
    with pytest.raises(gpu.NrsError):      # Helmholtz field without lambda1
        gpu.Elliptic(m, poisson=False, lambda0_field=l0, precon="jacobi")
