"""The two solver seams below ellipticSolve, through the C-ABI on the GPU.

* PRECONDITIONER = USER (elliptic_t::userPreconditioner, elliptic.h:178, called
  by ellipticPreconditioner.cpp:72-74): a hook that applies another handle's
  built-in multigrid reproduces that handle's solve bit for bit (same kernels,
  same order); a torch hook z = invDiagA r reproduces the built-in Jacobi solve
  (iterations +-1, solution 1e-8); errors raised in the hook and a missing hook
  fail loudly.
* pcg / pgmres alone (nrs_elliptic_krylov, PCG.cpp:85-203, PGMRES.cpp:111-340):
  given the residual that ellipticSolve forms (mask, then gather-scatter over the
  masked ids) and the solve's own tolerance, it takes the same iterations and
  returns the same solution as nrs_elliptic_solve from x0 = 0 (bit for bit; the
  bar written is 1e-13 relative)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPE, N = (4, 3, 3), 5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _rhs(m, seed):
    rng = np.random.default_rng(seed)
    return np.sin(np.pi * m.x) * np.cos(2 * np.pi * m.y) * (1 + m.z) + 1e-2 * rng.standard_normal(m.x.size)


MG = dict(precon="multigrid", levels=[5, 3, 1], smoother="chebyshev+ras", cheby_degree=2,
          coarse="smooth_jacobi", autotune=0)


@pytest.mark.parametrize("solver", ["fcg", "fgmres"])
def test_user_hook_gives_the_builtin_multigrid(gpu, torch_mod, solver):
    torch = torch_mod
    m = gpu.BoxMesh(*SHAPE, N, "kershaw", 0.3)
    a = gpu.Elliptic(m, solver=solver, tol=1e-10, max_iter=300, **MG)
    b = gpu.Elliptic(m, solver=solver, tol=1e-10, max_iter=300,
                     **dict(MG, precon="user"))
    lib = gpu._lib.load()
    calls = []

    def hook(r, z, stream):
        calls.append(1)
        rc = lib.nrs_elliptic_preconditioner(a._h, C.c_void_p(r), C.c_void_p(z),
                                             C.c_void_p(stream.cuda_stream))
        assert rc == 0

    b.set_user_preconditioner(hook, tensors=False)
    f = _rhs(m, 1)
    xa = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    xb = torch.zeros_like(xa)
    sa = a.solve(torch.from_numpy(f).cuda(), xa)
    sb = b.solve(torch.from_numpy(f).cuda(), xb)
    print(solver, "iterations builtin", sa["iterations"], "user", sb["iterations"], "hook calls",
          len(calls), flush=True)
    assert sa["converged"] and sb["converged"]
    assert sb["iterations"] == sa["iterations"]
    assert len(calls) >= sb["iterations"]
    assert torch.equal(xa, xb)


def test_user_hook_tensor_form_jacobi(gpu, torch_mod):
    """A torch hook z = invDiagA * r (the Jacobi handle's own diagonal) against the
    built-in Jacobi.  (Not against PRECONDITIONER = NONE: the reference's pcg then
    takes rdotz1 = the normalised residual norm, not r.r (PCG.cpp:129-130), so a
    scalar hook does not reproduce it; with this mesh NONE stalls.)"""
    torch = torch_mod
    m = gpu.BoxMesh(*SHAPE, N, "kershaw", 0.3)
    jac = gpu.Elliptic(m, solver="pcg", precon="jacobi", tol=1e-10, max_iter=2000, autotune=0)
    user = gpu.Elliptic(m, solver="pcg", precon="user", tol=1e-10, max_iter=2000, autotune=0)
    d = jac.buffer(4).clone()

    def hook(r, z, stream):
        assert r.dtype == torch.float64 and r.numel() == m.Nlocal and z.numel() == m.Nlocal
        with torch.cuda.stream(stream):
            torch.mul(d, r, out=z)

    user.set_user_preconditioner(hook)
    f = _rhs(m, 2)
    x0 = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    x1 = torch.zeros_like(x0)
    s0 = jac.solve(torch.from_numpy(f).cuda(), x0)
    s1 = user.solve(torch.from_numpy(f).cuda(), x1)
    print("iterations jacobi", s0["iterations"], "user", s1["iterations"],
          "diff", rel(x1.cpu().numpy(), x0.cpu().numpy()), flush=True)
    assert s0["converged"] and s1["converged"]
    assert abs(s0["iterations"] - s1["iterations"]) <= 1
    assert rel(x1.cpu().numpy(), x0.cpu().numpy()) <= 1e-8


def test_user_hook_errors(gpu, torch_mod):
    torch = torch_mod
    m = gpu.BoxMesh(2, 2, 2, 3)
    el = gpu.Elliptic(m, solver="pcg", precon="user", tol=1e-8, max_iter=5, autotune=0)
    f = torch.from_numpy(_rhs(m, 3)).cuda()
    x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    keep = f.clone()
    with pytest.raises(gpu.NrsError):            # no hook: rejected before any work
        el.solve(f, x)
    assert torch.equal(f, keep)

    def bad(r, z, stream):
        raise ValueError("planted")

    el.set_user_preconditioner(bad)
    with pytest.raises(RuntimeError) as e:
        el.solve(f.clone(), torch.zeros_like(x))
    assert isinstance(e.value.__cause__, ValueError)
    el.set_user_preconditioner(None)
    with pytest.raises(gpu.NrsError):
        el.solve(f.clone(), torch.zeros_like(x))


def _assembled_residual(gpu, torch, m, el, f):
    """mask(f), then gather-scatter over the ids with the Dirichlet nodes left out:
    the residual ellipticSolve forms from x0 = 0 (ellipticSolve.cpp:91-103)."""
    ids = np.array(m.gid, dtype=np.int64)
    mids = el.mask_ids.cpu().numpy()
    ids[mids] = 0
    og = gpu.Ogs(ids)
    r = torch.from_numpy(f).cuda()
    gpu.mask(el.mask_ids, r)
    og.gather_scatter(r)
    return r


@pytest.mark.parametrize("kw", [dict(solver="pcg", precon="jacobi"),
                                dict(MG, solver="fgmres", gmresRestart=6)])
def test_krylov_seam_matches_solve(gpu, torch_mod, kw):
    torch = torch_mod
    m = gpu.BoxMesh(*SHAPE, N, "kershaw", 0.3)
    kw = dict(kw)
    kw.setdefault("autotune", 0)
    el = gpu.Elliptic(m, tol=1e-9, max_iter=400, **kw)
    f = _rhs(m, 4)
    xs = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    st = el.solve(torch.from_numpy(f).cuda(), xs)
    r = _assembled_residual(gpu, torch, m, el, f)
    w = el.inv_degree.cpu().numpy()
    rh = r.cpu().numpy()
    res0_host = float(np.sqrt(np.sum(w * rh * rh) / el.volume))
    assert abs(res0_host - st["res0"]) <= 1e-13 * st["res0"]
    xk = torch.zeros_like(xs)
    it, res = el.krylov(r, xk, 1e-9 * st["res0"], 400, st["res0"])
    print(kw["solver"], "iterations solve", st["iterations"], "krylov", it, "res", st["res"], res,
          "bitwise", bool(torch.equal(xs, xk)), flush=True)
    assert st["converged"]
    assert it == st["iterations"]
    assert abs(res - st["res"]) <= 1e-13 * st["res0"]
    assert rel(xk.cpu().numpy(), xs.cpu().numpy()) <= 1e-13


def test_krylov_seam_max_iter_and_zero_residual(gpu, torch_mod):
    torch = torch_mod
    m = gpu.BoxMesh(*SHAPE, N, "kershaw", 0.3)
    el = gpu.Elliptic(m, solver="pcg", precon="jacobi", tol=1e-9, max_iter=400, autotune=0)
    r = _assembled_residual(gpu, torch, m, el, _rhs(m, 5))
    w = el.inv_degree.cpu().numpy()
    rh = r.cpu().numpy()
    res0 = float(np.sqrt(np.sum(w * rh * rh) / el.volume))
    x = torch.zeros_like(r)
    it, res = el.krylov(r, x, 1e-30, 7, res0)          # maxIter replaces opts.maxIter
    assert it == 7 and np.isfinite(res) and res > 0    # (CG's residual norm need not fall)
    assert float(x.abs().max()) > 0
    z = torch.zeros_like(r)
    x = torch.zeros_like(r)
    it, res = el.krylov(z, x, 1e-12, 50, 0.0)          # zero residual: pcg's do-while runs once
    assert it == 1 and res == 0.0 and float(x.abs().max()) == 0.0
    with pytest.raises(gpu.NrsError):
        el.krylov(r, x, 1e-9, 0, res0)


@pytest.mark.parametrize("tensors", [False, True])
def test_user_coarse_solve_gives_the_builtin_coarse_smoothing(gpu, torch_mod, tensors):
    """coarse="user" (coarseLevel_t::solvePtr, MG/MGSolver.hpp:99) with a hook that
    runs another handle's coarsest-level smoothing reproduces that handle's
    V-cycle and its FCG solve bit for bit; without a hook both fail before any work."""
    torch = torch_mod
    m = gpu.BoxMesh(*SHAPE, N, "kershaw", 0.3)
    kw = dict(precon="multigrid", solver="fcg", levels=[5, 3, 1], smoother="opt4chebyshev+ras",
              cheby_degree=2, autotune=0, tol=1e-10, max_iter=300)
    a = gpu.Elliptic(m, coarse="smooth_jacobi", **kw)
    b = gpu.Elliptic(m, coarse="user", **kw)
    nC = b.mg_info()[0][-1]["Nlocal"]
    r = torch.from_numpy(_rhs(m, 6)).cuda()
    zb = torch.zeros_like(r)
    with pytest.raises(gpu.NrsError):
        b.preconditioner(r, zb)
    with pytest.raises(gpu.NrsError):
        b.solve(r.clone(), torch.zeros_like(r))
    lib = gpu._lib.load()
    calls = []
    if tensors:
        def coarse(rhs, x, st):
            assert rhs.dtype == torch.float32 and rhs.numel() == nC and x.numel() == nC
            calls.append(1)
            a.mg_level(2, "smooth", rhs, x, x_is_zero=True, stream=st)
    else:
        def coarse(rhs, x, st):
            calls.append(1)
            rc = lib.nrs_elliptic_mg_level(a._h, 2, 2, C.c_void_p(rhs), None, C.c_void_p(x), 1,
                                           C.c_void_p(st.cuda_stream))
            assert rc == 0
    b.set_user_coarse_solver(coarse, tensors=tensors)
    za = torch.zeros_like(r)
    a.preconditioner(r, za)
    b.preconditioner(r, zb)
    torch.cuda.synchronize()
    assert len(calls) == 1 and torch.equal(za, zb)
    f = _rhs(m, 7)
    xa = torch.zeros_like(r)
    xb = torch.zeros_like(r)
    sa = a.solve(torch.from_numpy(f).cuda(), xa)
    sb = b.solve(torch.from_numpy(f).cuda(), xb)
    print("coarse hook: iterations", sa["iterations"], sb["iterations"], flush=True)
    assert sa["converged"] and sb["iterations"] == sa["iterations"] and torch.equal(xa, xb)
