"""GPU parity of the overlapping Schwarz / FDM smoother (ASM and RAS) and of
multigrid-preconditioned flexible CG with Chebyshev+Schwarz smoothing, against
the oracle (oracle/fdm.py, oracle/mg.py).  Bars: one smoother application
within 1e-12 relative in FP64 and 2e-5 in FP32 (different summation orders);
solves within +-1 iteration of the oracle and 1e-8 relative in the solution."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SM = {("cheby1", "asm"): "chebyshev+asm", ("opt4", "asm"): "opt4chebyshev+asm",
      ("cheby1", "ras"): "chebyshev+ras", ("opt4", "ras"): "opt4chebyshev+ras",
      (None, "asm"): "asm", (None, "ras"): "ras"}


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def pair(gpu, orc, shape, N, levels, cheby, S, prec=0, mp="kershaw", eps=0.3, degree=3, **kw):
    import oracle.mg as omg
    mg = gpu.BoxMesh(*shape, N, mp, eps)
    eg = gpu.Elliptic(mg, precon="multigrid", solver="fcg", levels=levels,
                      smoother=SM[(cheby, S)], cheby_degree=degree, mgPrecision=prec, **kw)
    mo = orc.BoxMesh(*shape, N, mp, eps)
    eo = orc.Elliptic(mo)
    T = np.float32 if prec == 0 else np.float64
    om = omg.MG(mo, levels, T=T, degree=degree, cheby=cheby, S=S)
    return eg, eo, om


@pytest.mark.parametrize("S", ["asm", "ras"])
@pytest.mark.parametrize("prec", [0, 1])
def test_level_smoother_vs_oracle(gpu, torch_mod, orc, S, prec):
    torch = torch_mod
    eg, eo, om = pair(gpu, orc, (4, 3, 3), 7, [7, 3, 1], "opt4", S, prec)
    T = np.float32 if prec == 0 else np.float64
    tdt = torch.float32 if prec == 0 else torch.float64
    for lvl in (0, 1):
        L = om.levels[lvl]
        x = np.random.default_rng(lvl).uniform(-1, 1, L.E * L.Np)
        x[L.el.maskIds] = 0
        x = L.el.ogs.gs(x, "add").astype(T)
        ref = L.schwarz.apply(x)
        out = torch.zeros(L.E * L.Np, dtype=tdt, device="cuda")
        eg.mg_smoother(lvl, dev(torch, x), out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        tol = 2e-5 if prec == 0 else 1e-12
        assert rel(got, ref) < tol, (lvl, rel(got, ref))
        assert np.all(got[L.el.maskIds] == 0)
    # lambda_max(S A) estimates agree
    lv, _ = eg.mg_info()
    for l, L in zip(lv[:-1], om.levels[:-1]):
        assert abs(l["rho"] - L.rho) <= 1e-4 * L.rho, (l["rho"], L.rho)


def _rotate(N, E, arrs, turns):
    Nq = N + 1
    out = []
    for a in arrs:
        b = a.reshape(E, Nq, Nq, Nq).copy()
        for e in range(E):
            b[e] = np.rot90(b[e], k=turns[e], axes=(1, 2))
        out.append(np.ascontiguousarray(b.ravel()))
    return out


def _rotate_etob(EToB, turns):
    out = EToB.reshape(-1, 6).copy()
    for e, t in enumerate(turns):
        for _ in range(t):      # one quarter turn: new faces (4, 2, 1, 3) <- old (1, 3, 2, 4)
            f = out[e].copy()
            out[e, 4], out[e, 2], out[e, 1], out[e, 3] = f[1], f[3], f[2], f[4]
    return np.ascontiguousarray(out.ravel())


@pytest.mark.parametrize("S", ["asm", "ras"])
def test_smoother_on_reoriented_elements(gpu, torch_mod, orc, S):
    """Elements stored in rotated local orientations: the face links must pair
    the same nodes the extended-mesh gather-scatter pairs."""
    import oracle.fdm as F
    torch = torch_mod
    N, shape = 5, (3, 3, 2)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    me = orc.BoxMesh(*shape, N + 2)
    turns = np.random.default_rng(3).integers(0, 4, mo.E)
    mo.x, mo.y, mo.z, mo.gid = _rotate(N, mo.E, (mo.x, mo.y, mo.z, mo.gid), turns)
    mo.EToB = _rotate_etob(mo.EToB, turns)
    (me.gid,) = _rotate(N + 2, me.E, (me.gid,), turns)
    me.EToB = _rotate_etob(me.EToB, turns)
    mo.ggeo = np.zeros(mo.E * 7 * mo.Np)
    orc.lib().orc_geofacs(N, mo.E, mo.D.ravel().copy(), mo.w, mo.x, mo.y, mo.z, mo.ggeo, None)
    eo = orc.Elliptic(mo)
    sw = F.Schwarz(mo, (N, mo.x, mo.y, mo.z, mo.w, eo.ogs_full), eo.maskIds, eo.ogs, eo.invDegree,
                   mode=S, T=np.float64, mesh_ext=me)
    eg = gpu.Elliptic(mo, precon="multigrid", solver="fcg", levels=[N, 1], mgPrecision=1,
                      smoother=SM[("opt4", S)], cheby_degree=3)
    x = np.random.default_rng(1).uniform(-1, 1, mo.Nlocal)
    x[eo.maskIds] = 0
    x = eo.ogs.gs(x, "add")
    ref = sw.apply(x)
    out = torch.zeros(mo.Nlocal, dtype=torch.float64, device="cuda")
    eg.mg_smoother(0, dev(torch, x), out)
    torch.cuda.synchronize()
    assert rel(out.cpu().numpy(), ref) < 1e-12


@pytest.mark.parametrize("cheby,S,degree", [("opt4", "asm", 3), ("opt4", "ras", 3),
                                            ("cheby1", "asm", 2), ("cheby1", "ras", 2)])
def test_mg_schwarz_fcg_vs_oracle(gpu, torch_mod, orc, cheby, S, degree):
    import oracle.mg as omg
    torch = torch_mod
    eg, eo, om = pair(gpu, orc, (4, 4, 4), 7, [7, 3, 1], cheby, S, degree=degree, tol=1e-10,
                      max_iter=200)
    rhs = np.random.default_rng(7).uniform(-1, 1, eg.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(cheby, S, degree, "gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8


@pytest.mark.parametrize("S", ["asm", "ras"])
def test_pure_schwarz_fcg_vs_oracle(gpu, torch_mod, orc, S):
    """ASM / RAS without Chebyshev (no post-smoothing) on a sine-deformed box."""
    import oracle.mg as omg
    torch = torch_mod
    eg, eo, om = pair(gpu, orc, (4, 4, 3), 7, [7, 3, 1], None, S, mp="sine", eps=0.0,
                      tol=1e-10, max_iter=200)
    rhs = np.random.default_rng(8).uniform(-1, 1, eg.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(S, "gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8


def test_schwarz_default_schedule(gpu, torch_mod):
    """Without explicit levels, ASM/RAS use the Schwarz schedule {7,3,1}
    (determineMGLevels.cpp:55-76)."""
    m = gpu.BoxMesh(3, 3, 3, 7, "kershaw", 0.3)
    el = gpu.Elliptic(m, precon="multigrid", solver="fcg", smoother="opt4chebyshev+ras",
                      cheby_degree=3)
    lv, _ = el.mg_info()
    assert [l["order"] for l in lv] == [7, 3, 1]


FMTS = {"bf16": 1, "fp16": 2}


def rel2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
@pytest.mark.parametrize("S", ["asm", "ras"])
def test_mfma_smoother_vs_oracle(gpu, torch_mod, orc, fmt, S):
    """BF16/FP16 contractions on the matrix cores against the oracle's
    emulation of the same roundings (FP32 accumulation), and against FP32."""
    import oracle.mg as omg
    torch = torch_mod
    mg = gpu.BoxMesh(4, 3, 3, 7, "kershaw", 0.3)
    eg = gpu.Elliptic(mg, precon="multigrid", solver="fcg", levels=[7, 3, 1],
                      smoother=SM[("opt4", S)], cheby_degree=3, fdmPrecision=FMTS[fmt])
    mo = orc.BoxMesh(4, 3, 3, 7, "kershaw", 0.3)
    om = omg.MG(mo, [7, 3, 1], T=np.float32, degree=3, cheby="opt4", S=S, fmt=fmt)
    o32 = omg.MG(mo, [7, 3, 1], T=np.float32, degree=3, cheby="opt4", S=S)
    for lvl in (0, 1):
        L = om.levels[lvl]
        x = np.random.default_rng(10 + lvl).uniform(-1, 1, L.E * L.Np)
        x[L.el.maskIds] = 0
        x = L.el.ogs.gs(x, "add").astype(np.float32)
        ref = L.schwarz.apply(x)
        ref32 = o32.levels[lvl].schwarz.apply(x)
        out = torch.zeros(L.E * L.Np, dtype=torch.float32, device="cuda")
        eg.mg_smoother(lvl, dev(torch, x), out)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        e_emul, e_f32 = rel2(got, ref), rel2(got, ref32)
        print(fmt, S, lvl, "vs emulation %.2e, vs FP32 %.2e, emulation vs FP32 %.2e" %
              (e_emul, e_f32, rel2(ref, ref32)))
        assert e_emul < (2e-3 if fmt == "bf16" else 3e-4)
        assert e_emul < 0.5 * rel2(ref, ref32)      # closer to the emulation than to FP32
        assert np.all(got[L.el.maskIds] == 0)


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_mg_mfma_fcg_vs_oracle(gpu, torch_mod, orc, fmt):
    import oracle.mg as omg
    torch = torch_mod
    mg = gpu.BoxMesh(4, 4, 4, 7, "kershaw", 0.3)
    eg = gpu.Elliptic(mg, precon="multigrid", solver="fcg", levels=[7, 3, 1], tol=1e-10,
                      max_iter=200, smoother="opt4chebyshev+ras", cheby_degree=3,
                      fdmPrecision=FMTS[fmt])
    mo = orc.BoxMesh(4, 4, 4, 7, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    om = omg.MG(mo, [7, 3, 1], T=np.float32, degree=3, cheby="opt4", S="ras", fmt=fmt)
    rhs = np.random.default_rng(7).uniform(-1, 1, eg.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros(eg.Nlocal, dtype=torch.float64, device="cuda")
    st = eg.solve(dev(torch, rhs), x)
    print(fmt, "gpu it", st["iterations"], "oracle it", info["iterations"])
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8
