"""GPU parity of the pressure right-hand side of the TOMBO splitting
(nrs_tombo_*, tombo::pressureSolve) against oracle/tombo.py.  Bars: geometry
and curl 1e-12, the full right-hand side 1e-11 relative to its largest entry
(FP64, different summation order), closed forms on an affine mesh to 1e-10."""
import types

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A = np.array([[1.3, 0.2, -0.1], [0.1, 0.9, 0.25], [-0.2, 0.15, 1.1]])
B = np.array([0.3, -0.2, 0.5])


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def meshes(gpu, orc, shape, N, mp="kershaw", eps=0.3, affine=False):
    mg = gpu.BoxMesh(*shape, N, mp, eps)
    mo = orc.BoxMesh(*shape, N, mp, eps)
    x, y, z = mo.x, mo.y, mo.z
    if affine:
        X = A @ np.stack([x, y, z]) + B[:, None]
        x, y, z = X[0].copy(), X[1].copy(), X[2].copy()
    m = types.SimpleNamespace(N=N, E=mo.E, x=x, y=y, z=z, gid=mg.gid, EToB=mg.EToB)
    return m, mo


def bc_mix(m, seed):
    """Boundary faces: inflow, wall, symmetry, outflow at random; 0 inside."""
    rng = np.random.default_rng(seed)
    kinds = np.array([3, 1, 7, 4, 16, 2])
    return np.where(m.EToB != 0, kinds[rng.integers(0, len(kinds), m.EToB.size)], 0).astype(np.int32)


@pytest.mark.parametrize("N", [1, 3, 7])
def test_geometry_and_inv_lmm(gpu, torch_mod, orc, N):
    import oracle.tombo as ot
    m, mo = meshes(gpu, orc, (3, 2, 2), N)
    t = gpu.PressureRhs(m, bc_mix(m, 0))
    g = ot.Geometry(N, m.x, m.y, m.z, m.gid)
    assert rel(t.vgeo.cpu().numpy().reshape(g.E, 10, -1), g.vgeo()) <= 1e-12
    assert rel(t.inv_lmm.cpu().numpy(), g.invLMM) <= 1e-13


@pytest.mark.parametrize("N", [2, 5, 7, 9])
def test_curl_vs_oracle(gpu, torch_mod, orc, N):
    import oracle.tombo as ot
    torch = torch_mod
    m, mo = meshes(gpu, orc, (3, 3, 2), N)
    t = gpu.PressureRhs(m)
    g = ot.Geometry(N, m.x, m.y, m.z, m.gid)
    L = m.x.size
    off = L + 37                                     # a padded field offset
    U = np.random.default_rng(N).uniform(-1, 1, 3 * off)
    W = t.curl(torch.from_numpy(U).cuda(), field_offset=off).cpu().numpy()
    ref = g.curl(U, off)
    got = np.concatenate([W[c * off:c * off + L] for c in range(3)])
    assert rel(got, ref) <= 1e-12


@pytest.mark.parametrize("N,visc,seed", [(7, True, 0), (7, False, 1), (4, True, 2), (1, True, 3)])
def test_pressure_rhs_vs_oracle(gpu, torch_mod, orc, N, visc, seed):
    import oracle.tombo as ot
    torch = torch_mod
    m, mo = meshes(gpu, orc, (4, 3, 3), N)
    bc = bc_mix(m, seed)
    t = gpu.PressureRhs(m, bc)
    g = ot.Geometry(N, m.x, m.y, m.z, m.gid)
    L = m.x.size
    off = L + 64
    rng = np.random.default_rng(seed)
    Ue, U, BF = (rng.uniform(-1, 1, 3 * off) for _ in range(3))
    mue = rng.uniform(0.5, 1.5, L)
    irho = rng.uniform(0.5, 1.5, L)
    div = rng.uniform(-0.1, 0.1, L)
    g0idt = 1.5 / 0.01
    ref = g.pressure_rhs(bc, Ue, U, BF, mue, irho, div, g0idt, visc, off)
    dv = lambda a: torch.from_numpy(a).cuda()
    got = t.compute(dv(Ue), dv(U), dv(BF), dv(mue), dv(irho), dv(div), g0idt, visc,
                    field_offset=off).cpu().numpy()
    print("N", N, "visc", visc, "flux elements", int(np.sum(np.any(np.isin(bc.reshape(-1, 6), [2, 3, 4, 7]), 1))),
          "rel", rel(got, ref), flush=True)
    assert rel(got, ref) <= 1e-11


@pytest.mark.parametrize("N,fusion", [(7, True), (7, False), (5, True)])
def test_pressure_rhs_stress_vs_oracle(gpu, torch_mod, orc, N, fusion):
    """VELOCITY STRESSFORMULATION = TRUE: the stress term of pressureStressHex3D
    taken from the second curl (oracle pinned by the closed form of
    test_oracle_tombo.py::test_stress_term_closed_form), with a varying viscosity;
    at N = 7 also through the face-fused curl kernels."""
    import oracle.tombo as ot
    torch = torch_mod
    m, mo = meshes(gpu, orc, (4, 3, 3), N)
    bc = bc_mix(m, 5)
    t = gpu.PressureRhs(m, bc, fusion=fusion, autotune=False, stress=True)
    g = ot.Geometry(N, m.x, m.y, m.z, m.gid)
    L = m.x.size
    off = L + 64
    rng = np.random.default_rng(7)
    Ue, U, BF = (rng.uniform(-1, 1, 3 * off) for _ in range(3))
    mue = 0.01 * (1.0 + 0.4 * np.sin(2 * m.x) * np.cos(m.y) + 0.2 * m.z)
    irho = rng.uniform(0.5, 1.5, L)
    div = rng.uniform(-0.1, 0.1, L)
    g0idt = 1.5 / 0.01
    ref = g.pressure_rhs(bc, Ue, U, BF, mue, irho, div, g0idt, True, off, stress=True)
    plain = g.pressure_rhs(bc, Ue, U, BF, mue, irho, div, g0idt, True, off, stress=False)
    dv = lambda a: torch.from_numpy(a).cuda()
    got = t.compute(dv(Ue), dv(U), dv(BF), dv(mue), dv(irho), dv(div), g0idt, True,
                    field_offset=off).cpu().numpy()
    print("stress N", N, "fusion", fusion, "rel", rel(got, ref), "term size", rel(ref, plain), flush=True)
    assert rel(ref, plain) > 1e-8                  # the term is not negligible here
    assert rel(got, ref) <= 1e-11


def test_pressure_rhs_closed_form_affine(gpu, torch_mod, orc):
    """Walls everywhere: sum_i x_i rhs_i = sum JW (x + nu (2 + 4/3 a_x)) + g0idt sum x JW div
    for Ue = (y^2, z^2, x^2), BF = JW (x, 1, z), div = c + a.x (the oracle's closed form)."""
    import oracle.tombo as ot
    torch = torch_mod
    N = 4
    m, mo = meshes(gpu, orc, (3, 2, 2), N, mp="none", eps=0.0, affine=True)
    bc = np.where(m.EToB != 0, 1, 0).astype(np.int32)
    t = gpu.PressureRhs(m, bc)
    x, y, z = m.x, m.y, m.z
    L = x.size
    JW = ot.Geometry(N, x, y, z, m.gid).JW.ravel()
    nu = 0.7 * 1.5
    div = 0.3 + 0.5 * x - 0.2 * y
    dv = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    rhs = t.compute(dv(np.concatenate([y * y, z * z, x * x])), dv(np.zeros(3 * L)),
                    dv(np.concatenate([JW * x, JW, JW * z])), dv(np.full(L, 0.7)),
                    dv(np.full(L, 1.5)), dv(div), 4.0, True).cpu().numpy()
    expect = np.sum(JW * (x + nu * (2 + 4.0 / 3.0 * 0.5))) + 4.0 * np.sum(x * JW * div)
    assert abs(np.sum(x * rhs) - expect) <= 1e-10 * abs(expect)


def test_pressure_rhs_rejects_short_offset(gpu, torch_mod, orc):
    torch = torch_mod
    m, mo = meshes(gpu, orc, (2, 2, 2), 3)
    t = gpu.PressureRhs(m)
    L = m.x.size
    a = torch.zeros(3 * L, dtype=torch.float64, device="cuda")
    s = torch.zeros(L, dtype=torch.float64, device="cuda")
    with pytest.raises(gpu.NrsError):
        t.compute(a, a, a, s, s, s, 1.0, True, field_offset=L - 1)


def test_load_schedules_agree(gpu, torch_mod, orc):
    """Both load schedules of the curl kernels (NRS_TUNE_TOMBO 1 and 2) give
    the same right-hand side to rounding (the compiler may contract the
    multiply-adds of the two schedules differently): 400 eps of the largest entry."""
    torch = torch_mod
    N = 7
    m, mo = meshes(gpu, orc, (4, 3, 3), N)
    t = gpu.PressureRhs(m, bc_mix(m, 5))
    assert gpu.tuned_variant("tombo", N, "curl", m.E) in (1, 2)
    assert gpu.tuned_variant("tombo", N, "rhs", m.E) in (1, 2)
    L = m.x.size
    rng = np.random.default_rng(5)
    dv = lambda a: torch.from_numpy(a).cuda()
    args = [dv(rng.uniform(-1, 1, 3 * L)) for _ in range(3)] + [dv(rng.uniform(0.5, 1.5, L)) for _ in range(2)]
    args.append(dv(rng.uniform(-0.1, 0.1, L)))
    outs = []
    for v in (1, 2):
        gpu.set_variant("tombo", N, "curl", m.E, v)
        gpu.set_variant("tombo", N, "rhs", m.E, v)
        outs.append(t.compute(*args, 75.0, True).cpu().numpy())
    gpu.set_variant("tombo", N, "curl", m.E, 0)
    gpu.set_variant("tombo", N, "rhs", m.E, 0)
    print("schedules differ by", rel(outs[1], outs[0]), flush=True)
    assert rel(outs[1], outs[0]) <= 400 * np.finfo(np.float64).eps


@pytest.mark.parametrize("visc", [True, False])
def test_face_fusion_bitwise(gpu, torch_mod, orc, visc):
    """The curl kernels' face fusion (N = 7, consecutive elements of a block of four)
    gives the unfused right-hand side bit for bit, and matches the oracle."""
    import oracle.tombo as ot
    torch = torch_mod
    N = 7
    m, mo = meshes(gpu, orc, (8, 3, 2), N)
    bc = bc_mix(m, 9)
    on = gpu.PressureRhs(m, bc, fusion=True, autotune=False)
    off = gpu.PressureRhs(m, bc, fusion=False, autotune=False)
    L = m.x.size
    off_ = L + 16
    rng = np.random.default_rng(9)
    Ue, U, BF = (rng.uniform(-1, 1, 3 * off_) for _ in range(3))
    mue, irho = rng.uniform(0.5, 1.5, L), rng.uniform(0.5, 1.5, L)
    div = rng.uniform(-0.1, 0.1, L)
    dv = lambda a: torch.from_numpy(a).cuda()
    args = (dv(Ue), dv(U), dv(BF), dv(mue), dv(irho), dv(div), 80.0, visc)
    a = on.compute(*args, field_offset=off_).cpu().numpy()
    b = off.compute(*args, field_offset=off_).cpu().numpy()
    assert np.array_equal(a, b)
    g = ot.Geometry(N, m.x, m.y, m.z, m.gid)
    assert rel(a, g.pressure_rhs(bc, Ue, U, BF, mue, irho, div, 80.0, visc, off_)) <= 1e-11


def _rhs_inputs(torch, L, off, seed):
    rng = np.random.default_rng(seed)
    dv = lambda a: torch.from_numpy(a).cuda()
    Ue, U, BF = (dv(rng.uniform(-1, 1, 3 * off)) for _ in range(3))
    return (Ue, U, BF, dv(rng.uniform(0.5, 1.5, L)), dv(rng.uniform(0.5, 1.5, L)),
            dv(rng.uniform(-0.1, 0.1, L)))


@pytest.mark.parametrize("N", [1, 3, 5, 7, 9])
def test_corner_geometry_matches_stored(gpu, torch_mod, orc, N):
    """On trilinear elements (the Kershaw box) the volume kernels form rx..tz, JW
    from the element corners; the right-hand side, the curl and the stress term
    agree with the kernels that stream the stored factors to 1e-13 relative (the
    formed factors differ from the stored ones by rounding only)."""
    torch = torch_mod
    m, mo = meshes(gpu, orc, (4, 3, 3), N)
    bc = bc_mix(m, N)
    L = m.x.size
    off = L + 32
    for stress in (False, True):
        on = gpu.PressureRhs(m, bc, stress=stress)
        off_h = gpu.PressureRhs(m, bc, stress=stress, stored_geometry=True)
        assert on.corner_geometry and not off_h.corner_geometry
        args = _rhs_inputs(torch, L, off, N)
        for visc in ((True, False) if not stress else (True,)):
            a = on.compute(*args, 90.0, visc, field_offset=off).cpu().numpy()
            b = off_h.compute(*args, 90.0, visc, field_offset=off).cpu().numpy()
            print("N", N, "stress", stress, "visc", visc, "rel", rel(a, b), flush=True)
            assert rel(a, b) <= 1e-13
        wa = on.curl(args[1], field_offset=off).cpu().numpy()
        wb = off_h.curl(args[1], field_offset=off).cpu().numpy()
        idx = np.concatenate([np.arange(c * off, c * off + L) for c in range(3)])
        assert rel(wa[idx], wb[idx]) <= 1e-13


def test_corner_geometry_off_on_curved_mesh(gpu, torch_mod, orc):
    """Sine-mapped nodes are not trilinear: the kernels stream the stored
    factors whatever the flag says, bit for bit."""
    torch = torch_mod
    N = 5
    m = gpu.BoxMesh(3, 3, 2, N, "sine", 0.0)
    bc = bc_mix(m, 3)
    a_h = gpu.PressureRhs(m, bc)
    b_h = gpu.PressureRhs(m, bc, stored_geometry=True)
    assert not a_h.corner_geometry
    L = m.x.size
    args = _rhs_inputs(torch, L, L, 4)
    a = a_h.compute(*args, 70.0, True).cpu().numpy()
    b = b_h.compute(*args, 70.0, True).cpu().numpy()
    assert np.array_equal(a, b)
