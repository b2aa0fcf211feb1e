"""Kernel launch variants and the autotuner (KERNEL AUTOTUNING; the reference
times its p_knl variants at setup and drops any whose result differs from the
reference kernel's: src/bench/axHelm/benchmarkAx.cpp:268-383).

Every Ax variant (elements per block x1, x2, x4) meets the oracle bar of the
default launch (FP64 1e-12, FP32 2e-5 relative), also through an element
list; every matrix-core FDM variant (elements per wavefront 1, 2, 4, 8) gives
the default variant's smoother output bit for bit; a setup with autotuning
records a valid variant for each level's key and solves to the oracle's
iteration count."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("N", [1, 2, 3, 5, 7, 9])
@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("mult", [1, 2, 4])
def test_ax_variants_vs_oracle(gpu, torch_mod, orc, N, T, mult):
    torch = torch_mod
    rng = np.random.default_rng(7 * N + mult)
    E = 53                                  # not a multiple of any packing
    m = orc.BoxMesh(E, 1, 1, N, "kershaw", 0.3)
    Np = m.Np
    q = rng.uniform(-1, 1, E * Np)
    tname = "f64" if T == np.float64 else "f32"
    tt = torch.float64 if T == np.float64 else torch.float32
    tol = 1e-12 if T == np.float64 else 2e-5
    l0 = rng.uniform(0.5, 2, E * Np)
    l1 = rng.uniform(0, 10, E * Np)
    ref = orc.ax(N, m.ggeo.astype(T), m.D.astype(T), q.astype(T), lambda0=l0.astype(T),
                 lambda1=l1.astype(T), poisson=False, dtype=T)
    # a ragged element list: every third element, reversed
    elist = np.arange(E, dtype=np.int32)[::3][::-1].copy()
    gpu.set_variant("ax", N, tname, E, mult)
    gpu.set_variant("ax", N, tname, len(elist), mult)
    try:
        assert gpu.tuned_variant("ax", N, tname, E) == mult
        Aq = torch.zeros(E * Np, dtype=tt, device="cuda")
        gpu.ax(N, dev(torch, m.ggeo.astype(T)), dev(torch, m.D.astype(T).ravel()),
               dev(torch, q.astype(T)), Aq, lambda0=dev(torch, l0.astype(T)),
               lambda1=dev(torch, l1.astype(T)), poisson=False, lambda_field=True)
        torch.cuda.synchronize()
        assert rel(Aq.cpu().numpy(), ref) < tol
        Al = torch.full((E * Np,), 7.0, dtype=tt, device="cuda")
        gpu.ax(N, dev(torch, m.ggeo.astype(T)), dev(torch, m.D.astype(T).ravel()),
               dev(torch, q.astype(T)), Al, lambda0=dev(torch, l0.astype(T)),
               lambda1=dev(torch, l1.astype(T)), poisson=False, lambda_field=True,
               element_list=dev(torch, elist))
        torch.cuda.synchronize()
        got = Al.cpu().numpy().reshape(E, Np)
        listed = np.zeros(E, dtype=bool)
        listed[elist] = True
        assert rel(got[listed], ref.reshape(E, Np)[listed]) < tol
        assert np.all(got[~listed] == 7.0)
    finally:
        gpu.set_variant("ax", N, tname, E, 0)
        gpu.set_variant("ax", N, tname, len(elist), 0)


def test_set_variant_rejects_unknown(gpu):
    with pytest.raises(gpu.NrsError):
        gpu.set_variant("ax", 3, "f64", 100, 3)
    with pytest.raises(gpu.NrsError):
        gpu.set_variant("fdm", 3, "bf16", 100, 16)
    with pytest.raises(gpu.NrsError):
        gpu.set_variant("ax", 10, "f64", 100, 1)


@pytest.mark.parametrize("smoother,fmt", [("opt4chebyshev+ras", "bf16"), ("opt4chebyshev+asm", "fp16"),
                                          ("ras", "fp16")])
def test_fdm_variants_bitwise(gpu, torch_mod, smoother, fmt):
    torch = torch_mod
    m = gpu.BoxMesh(5, 4, 3, 7, "kershaw", 0.2)
    E = m.E
    el = gpu.Elliptic(m, solver="fcg", precon="multigrid", smoother=smoother, levels=[7, 3, 1],
                      fdmPrecision=gpu.FDM_PRECISIONS[fmt], autotune=0)
    x = torch.from_numpy(np.random.default_rng(3).uniform(-1, 1, m.Nlocal).astype(np.float32)).cuda()
    for lvl, N in ((0, 7), (1, 3)):
        L = E * (N + 1) ** 3
        xl = x[:L].contiguous()
        base = torch.zeros(L, dtype=torch.float32, device="cuda")
        el.mg_smoother(lvl, xl, base)
        torch.cuda.synchronize()
        assert torch.isfinite(base).all()
        try:
            for epw in (1, 2, 4, 8):
                gpu.set_variant("fdm", N, fmt, E, epw)
                out = torch.full((L,), float("nan"), dtype=torch.float32, device="cuda")
                el.mg_smoother(lvl, xl, out)
                torch.cuda.synchronize()
                assert torch.equal(out, base), (lvl, epw)
        finally:
            gpu.set_variant("fdm", N, fmt, E, 0)


def test_autotune_setup_records_and_solves(gpu, torch_mod, orc):
    import oracle.mg as omg
    torch = torch_mod
    shape, N = (4, 3, 3), 5
    m = gpu.BoxMesh(*shape, N, "kershaw", 0.25)
    kw = dict(solver="fcg", precon="multigrid", smoother="opt4chebyshev+ras", cheby_degree=3,
              levels=[N, 3, 1], tol=1e-10, max_iter=200, fdmPrecision=1)
    el = gpu.Elliptic(m, autotune=1, **kw)
    E = m.E
    assert gpu.tuned_variant("ax", N, "f64", E) in (1, 2, 4)
    for n in (N, 3, 1):
        assert gpu.tuned_variant("ax", n, "f32", E) in (1, 2, 4)
    for n in (N, 3):
        assert gpu.tuned_variant("fdm", n, "bf16", E) in (1, 2, 4, 8)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.25)
    eo = orc.Elliptic(mo)
    om = omg.MG(mo, [N, 3, 1], T=np.float32, degree=3, cheby="opt4", S="ras", fmt="bf16")
    rhs = np.random.default_rng(11).uniform(-1, 1, mo.Nlocal)
    xo, info = omg.fcg_solve(eo, rhs, om.apply, flexible=True, tol=1e-10, maxit=200)
    x = torch.zeros(m.Nlocal, dtype=torch.float64, device="cuda")
    st = el.solve(dev(torch, rhs), x)
    assert st["converged"]
    assert abs(st["iterations"] - info["iterations"]) <= 1
    assert rel(x.cpu().numpy(), xo) <= 1e-8
    assert gpu.tuned_variant("op", N, "f64", E) in (1, 2, 4, 8, 16)
    assert gpu.tuned_variant("op", N, "f32", E) in (1, 2, 4, 8, 16)
    for key in (("ax", N, "f64"), ("ax", N, "f32"), ("ax", 3, "f32"), ("ax", 1, "f32"),
                ("fdm", N, "bf16"), ("fdm", 3, "bf16"), ("op", N, "f64"), ("op", N, "f32")):
        gpu.set_variant(*key, E, 0)


@pytest.mark.parametrize("shape,N", [((8, 7, 6), 7), ((5, 1, 1), 3), ((16, 16, 9), 5)])
def test_operator_chunks_bitwise_and_vs_oracle(gpu, torch_mod, orc, shape, N):
    """The chunked operator (Ax of chunk c+1 overlapped with the gather-scatter
    of the rows chunk c completes) gives the unchunked result bit for bit for
    every chunk count, including meshes with fewer elements than chunks, and
    matches the oracle's operator (FP64 1e-12); the FP32 multigrid V-cycle,
    whose finest-level operator is chunked the same way, is bitwise unchanged."""
    torch = torch_mod
    m = gpu.BoxMesh(*shape, N, "kershaw", 0.3)
    E = m.E
    el = gpu.Elliptic(m, solver="fcg", precon="multigrid", smoother="opt4chebyshev+jacobi",
                      levels=[N, 1], autotune=0)
    mo = orc.BoxMesh(*shape, N, "kershaw", 0.3)
    eo = orc.Elliptic(mo)
    q = np.random.default_rng(N).uniform(-1, 1, m.Nlocal)
    ref = eo.operator(q)
    qd = dev(torch, q)
    outs, sm = [], []
    try:
        for k in (1, 2, 4, 8, 16):
            gpu.set_variant("op", N, "f64", E, k)
            gpu.set_variant("op", N, "f32", E, k)
            Aq = torch.full_like(qd, float("nan"))
            el.operator(qd, Aq)
            z = torch.full_like(qd, float("nan"))
            el.preconditioner(qd, z)
            torch.cuda.synchronize()
            outs.append(Aq)
            sm.append(z)
    finally:
        gpu.set_variant("op", N, "f64", E, 0)
        gpu.set_variant("op", N, "f32", E, 0)
    assert rel(outs[0].cpu().numpy(), ref) < 1e-12
    for k in range(1, len(outs)):
        assert torch.equal(outs[k], outs[0]), k
        assert torch.equal(sm[k], sm[0]), k
