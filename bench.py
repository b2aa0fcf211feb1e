#!/usr/bin/env python3
"""Benchmark: FP64 Ax + gather-scatter throughput on one MI355X.

Workload (the nekRS C1 shape): unit cube, 32x32x32 = 32768 hexahedra at N=7
(16.8M local nodes, 11.4M unique), Kershaw eps=0.3 trilinear elements, six
geometric factors per node precomputed, q ~ U(-1,1) (seed 0).  One step is
one ellipticOperator application Aq = gs(mask(Ax(q))) in FP64; the default run
times 100 steps after 10 warm-up steps, inputs resident in HBM.

  value      GDOF/s = E*N^3 / (seconds per step) / 1e9   (reference convention,
             src/bench/axHelm/benchmarkAx.cpp:312); the operator as set up by
             default, i.e. with the face fusion (the interior nodes of faces
             shared by consecutive elements of a block of four summed inside
             the Ax kernel, bit-identical results); the same loop without it
             is timed on the same box and reported in config
  roofline   dominant kernel = FP64 Ax: 64 algorithmic bytes per local node
             (q 8 + Aq 8 + 6 geometric factors 48) / average launch time
             measured with HIP events around every Ax launch in the timed loop
  cpu_baseline  the operator on the host cores on the same mesh: the reference's
             own serial CPU kernel and gslib where oracle/_ref is built, the
             oracle's OpenMP port beside it
  solves     the second half of the metric: time to a 1e-8 relative residual of
             the 48^3 N=7 Kershaw pressure-type Poisson solve (manufactured sine
             solution plus seeded 1e-3 Gaussian noise in the forcing), flexible
             CG preconditioned by p-multigrid 7-3-1 with an FP64 coarse solve (two
             PCG iterations on the assembled order-1 matrix, each preconditioned by
             one smoothed-aggregation AMG V-cycle, on the device; the V-cycle is
             replayed as one HIP graph): all-FP64 Chebyshev-Jacobi (degree 2)
             against FP32 multigrid with optimal 4th-kind Chebyshev + RAS/FDM
             smoothing (degree 4), the FDM contractions in BF16 on the matrix cores
             (FP32 accumulation).  Each case is set up once, solved once untimed,
             then timed over one solve from x = 0 (setup excluded).
  pressure_rhs  the step before the solve: tombo::pressureSolve's right-hand
             side on the same 32^3 mesh (channel boundary faces), ms per call and
             algorithmic GB/s, beside the same call streaming stored factors
  plain_poisson  the reference's CPU-runnable case: 8^3 N=7 sine-deformed cube,
             Jacobi PCG to 1e-8, FP64, and the oracle's solve on the host beside it
  kershaw    the reference's own published benchmarks (examples/kershaw): BPS5,
             the pressure solve with the reference's defaults (PGMRES+FLEXIBLE,
             RAS + opt. 4th-kind Chebyshev degree 3, FP32 multigrid) on the 20^3
             N=7 Kershaw box, with the FDM in FP32 and in BF16, and BP5
             (unpreconditioned PCG, 1000 iterations), beside the published
             V100 / A100 / MI250X figures
  batch      64 uneven Helmholtz problems advanced by captured graphs, and the
             3-component velocity Helmholtz solve (batch of 3, and one block solve)

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "Ax+gs GDOF/s (FP64); s to 1e-8 rel-res, mixed vs FP64 p-MG, 1 MI355X"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


class Ranges:
    """roctx ranges around each timed solve, so that a `rocprofv3 --marker-trace`
    run can attribute kernels to them; no-ops when the roctx library is absent."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                self.lib = ctypes.CDLL(name)
                self.lib.roctxRangePushA.argtypes = [ctypes.c_char_p]
                break
            except (OSError, AttributeError):
                self.lib = None

    def push(self, name):
        if self.lib is not None:
            self.lib.roctxRangePushA(name.encode())

    def pop(self):
        if self.lib is not None:
            self.lib.roctxRangePop()


RANGES = None


def ranges():
    global RANGES
    if RANGES is None:
        RANGES = Ranges()
    return RANGES


def traffic_from_profiles():
    """HBM bytes per Ax launch from the committed rocprofv3 --pmc summary."""
    p = os.path.join(ROOT, "profiles", "ax_f64_n7_e32768_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get("bytes_per_launch")
    except Exception:
        return None


def cpu_baseline(nx, N, seconds, seed):
    """The operator (Ax + mask + gather-scatter, FP64) on the host cores, on the
    same mesh, over a bounded sample.  Where oracle/_ref/libnrsref.so was built
    (from the reference's own sources, by oracle/build_ref.sh), the timed path is
    the reference's: its serial CPU kernel ellipticPartialAxCoeffHex3D.c (OpenMP
    over elements) and its vendored gslib gs (serial); the oracle's OpenMP
    operator is timed beside it ("port").  Otherwise the port alone."""
    import ctypes as C
    import numpy as np
    import oracle
    oracle.build()
    t0 = time.perf_counter()
    m = oracle.BoxMesh(nx, nx, nx, N, "kershaw", 0.3)
    el = oracle.Elliptic(m)
    q = np.random.default_rng(seed).uniform(-1, 1, m.Nlocal)
    log("cpu baseline setup %.1f s" % (time.perf_counter() - t0))
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)

    def timed(fn, budget):
        fn()                                # warm-up
        reps, t = 0, 0.0
        while t < budget or reps < 2:
            a = time.perf_counter()
            fn()
            t += time.perf_counter() - a
            reps += 1
        return reps, t

    ref = oracle.ref_lib()
    port_budget = seconds if ref is None else seconds / 3
    reps, t = timed(lambda: el.operator(q), port_budget)
    port = {"value": m.E * N ** 3 * reps / t / 1e9, "unit": "GDOF/s", "cores": cores, "kind": "port",
            "sample": "%d operator applications (Ax+mask+gs, FP64) on the same %d^3 N=%d "
                      "Kershaw mesh, %.1f s" % (reps, nx, N, t)}
    if ref is None:
        return port
    Nq = N + 1
    ax = getattr(ref, "ref_ax_poisson_double_nq%d" % Nq)
    dp = lambda a: a.ctypes.data_as(C.c_void_p)
    E, L = m.E, m.Nlocal
    elist = np.arange(E, dtype=np.int32)
    D = np.ascontiguousarray(m.D, dtype=np.float64)
    S = np.ascontiguousarray(D.T)
    l0, l1 = np.ones(1), np.zeros(1)
    gid = oracle.masked_gids(m.gid, el.maskIds)
    ref.ref_gslib_setup.restype = C.c_void_p
    h = C.c_void_p(ref.ref_gslib_setup(dp(gid), C.c_uint(L)))
    Aq = np.zeros(L)
    mask = el.maskIds
    cE, cL = C.c_int(E), C.c_int(L)          # the kernel takes its sizes by reference (dlong &)

    def apply():
        ax(C.byref(cE), C.byref(cL), C.byref(cL), dp(elist), dp(m.ggeo), dp(D), dp(S), dp(l0), dp(l1),
           dp(q), dp(Aq))
        Aq[mask] = 0.0
        ref.ref_gslib_apply(h, dp(Aq), 0, 0)
    try:
        apply()
        check = el.operator(q)
        err = float(np.max(np.abs(Aq - check)) / np.max(np.abs(check)))
        reps, t = timed(apply, seconds * 2 / 3)
    finally:
        ref.ref_gslib_free(h)
    return {"value": E * N ** 3 * reps / t / 1e9, "unit": "GDOF/s", "cores": cores, "kind": "reference",
            "sample": "%d operator applications (the reference's serial CPU kernel "
                      "ellipticPartialAxCoeffHex3D.c, OpenMP over elements, built from its sources; "
                      "mask; its gslib gs, serial) on the same %d^3 N=%d Kershaw mesh, %.1f s; "
                      "rel. difference to the oracle's operator %.1e" % (reps, nx, N, t, err),
            "port": port, "port_over_reference": port["value"] / (E * N ** 3 * reps / t / 1e9)}


def solve_cases(n, N, seed=0, only=None):
    """The FP64 and mixed-precision p-MG solves on the n^3 Kershaw mesh (`only`:
    the names of the cases to run, all by default)."""
    import numpy as np
    import torch
    import nekrs_amd
    mesh = nekrs_amd.BoxMesh(n, n, n, N, "kershaw", 0.3)
    u = np.sin(np.pi * mesh.x) * np.sin(np.pi * mesh.y) * np.sin(np.pi * mesh.z)
    f = 3 * np.pi ** 2 * u + 1e-3 * np.random.default_rng(seed).standard_normal(u.size)
    amg = dict(coarse="amg", coarse_tol=0.0, coarse_max_iter=2)
    cases = [("fp64_chebyshev_jacobi", dict(smoother="chebyshev+jacobi", cheby_degree=2,
                                            mgPrecision=1, **amg)),
             ("mixed_opt4chebyshev_ras", dict(smoother="opt4chebyshev+ras", cheby_degree=4,
                                              mgPrecision=0, fdmPrecision=1, **amg)),
             # the reference's default outer solver (PGMRES+FLEXIBLE, restart 15) on the
             # mixed preconditioner
             ("mixed_fgmres", dict(smoother="opt4chebyshev+ras", cheby_degree=4, mgPrecision=0,
                                   fdmPrecision=1, solver="fgmres", **amg))]
    if only:
        cases = [c for c in cases if c[0] in only]
    out, xs = {}, {}
    for name, kw in cases:
        t0 = time.perf_counter()
        kw = dict(kw)
        solver = kw.pop("solver", "fcg")
        el = nekrs_amd.Elliptic(mesh, precon="multigrid", solver=solver, tol=1e-8, max_iter=1000,
                                levels=[N, 3, 1], **kw)
        torch.cuda.synchronize()
        t_setup = time.perf_counter() - t0
        Np = (N + 1) ** 3
        gw = el.ggeo.cpu().numpy().reshape(-1, 7, Np)[:, 6].ravel()
        rhs = torch.from_numpy(gw * f).cuda()
        for rep in range(2):
            r = rhs.clone()
            x = torch.zeros_like(r)
            torch.cuda.synchronize()
            if rep == 1:
                ranges().push("nrs:solve:" + name)
            st = el.solve(r, x)
            if rep == 1:
                ranges().pop()
        # true residual of the returned solution, in the solver's norm
        b = rhs.clone()
        mk = el.mask_ids.long()
        b[mk] = 0
        mgid = mesh.gid.copy()
        mgid[mk.cpu().numpy()] = 0
        nekrs_amd.Ogs(mgid).gather_scatter(b, "add")
        Ax = torch.empty_like(x)
        el.operator(x, Ax)
        w = el.inv_degree
        res_true = torch.sqrt(torch.sum(w * (b - Ax) ** 2) / el.volume).item()
        xh = x.cpu().numpy()
        xs[name] = xh
        out[name] = {"seconds": st["seconds"], "iterations": st["iterations"],
                     "converged": st["converged"], "setup_seconds": t_setup,
                     "true_rel_res": res_true / st["res0"],
                     "err_vs_exact_max": float(np.max(np.abs(xh - u)))}
        log("%s: %d it, %.3f s (setup %.1f s), true rel res %.2e" %
            (name, st["iterations"], st["seconds"], t_setup, res_true / st["res0"]))
        del el
    if "fp64_chebyshev_jacobi" in xs and "mixed_opt4chebyshev_ras" in xs:
        a, b_ = xs["fp64_chebyshev_jacobi"], xs["mixed_opt4chebyshev_ras"]
        out["mixed_opt4chebyshev_ras"]["err_vs_fp64_rel"] = float(np.max(np.abs(b_ - a)) /
                                                                  np.max(np.abs(a)))
        out["speedup_mixed_vs_fp64"] = (out["fp64_chebyshev_jacobi"]["seconds"] /
                                        out["mixed_opt4chebyshev_ras"]["seconds"])
    out["workload"] = ("%d^3=%d hex elements N=%d Kershaw eps=0.3, p-MG %d-3-1 with an FP64 coarse "
                       "solve (2 AMG-preconditioned PCG iterations), FCG to 1e-8 relative, rhs "
                       "sine + 1e-3 noise (seed %d)" % (n, n ** 3, N, N, seed))
    return out


def plain_poisson_case(seed=0, cpu=False):
    """The reference's CPU-runnable case: 8^3 = 512 hex elements at N=7 on the unit
    cube, nodes moved by 0.05 sin(pi x) sin(pi y) sin(pi z), manufactured solution
    sin(pi x) sin(pi y) sin(pi z) with Dirichlet boundaries, Jacobi PCG to 1e-8
    relative, FP64 throughout; with cpu, the oracle's Jacobi PCG on the host cores
    beside it (same mesh and right-hand side)."""
    import numpy as np
    import torch
    import nekrs_amd
    mesh = nekrs_amd.BoxMesh(8, 8, 8, 7, "sine", 0.0)
    el = nekrs_amd.Elliptic(mesh, precon="jacobi", solver="pcg", tol=1e-8, max_iter=1000)
    Np = 8 ** 3
    u = np.sin(np.pi * mesh.x) * np.sin(np.pi * mesh.y) * np.sin(np.pi * mesh.z)
    gw = el.ggeo.cpu().numpy().reshape(-1, 7, Np)[:, 6].ravel()
    rhs = torch.from_numpy(3 * np.pi ** 2 * u * gw).cuda()
    st = None
    for rep in range(3):
        x = torch.zeros_like(rhs)
        torch.cuda.synchronize()
        st = el.solve(rhs.clone(), x)
    err = float(np.max(np.abs(x.cpu().numpy() - u)))
    out = {"iterations": st["iterations"], "seconds": st["seconds"], "converged": st["converged"],
           "err_vs_exact_max": err,
           "workload": "8^3=512 hex elements N=7, sine-deformed unit cube, manufactured sine "
                       "solution, Dirichlet, Jacobi PCG to 1e-8 relative, FP64"}
    log("plain Poisson 8^3 N=7 Jacobi PCG: %d it, %.4f s, max error vs exact %.2e" %
        (st["iterations"], st["seconds"], err))
    if cpu:
        import oracle
        oracle.build()
        mo = oracle.BoxMesh(8, 8, 8, 7, "sine", 0.0)
        eo = oracle.Elliptic(mo)
        b = 3 * np.pi ** 2 * u * mo.ggeo.reshape(-1, 7, Np)[:, 6].ravel()
        eo.solve(b, tol=1e-8, maxit=1000)                 # warm-up
        t0 = time.perf_counter()
        _, info = eo.solve(b, tol=1e-8, maxit=1000)
        out["cpu_port"] = {"seconds": time.perf_counter() - t0, "iterations": info["iterations"],
                           "cores": int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)}
        log("plain Poisson on the host (oracle, %d cores): %d it, %.3f s" %
            (out["cpu_port"]["cores"], info["iterations"], out["cpu_port"]["seconds"]))
    return out


# the reference's published Kershaw results at 8000 elements per GPU, N=7
# (examples/kershaw/README.md:13-57): solve seconds / iterations for BPS5 and
# (DOF x iter)/s for BP5
PUBLISHED_KERSHAW = {"V100": {"bps5_s": 0.292781, "bps5_it": 28, "bp5_dof_it_per_s": 2.39e9},
                     "A100": {"bps5_s": 0.196772, "bps5_it": 28, "bp5_dof_it_per_s": 3.87e9},
                     "MI250X/1": {"bps5_s": 0.279318, "bps5_it": 28, "bp5_dof_it_per_s": 3.07e9}}


def kershaw_cases(N=7, reps=5):
    """The reference's Kershaw benchmarks (examples/kershaw/kershaw.udf:14-140) on
    its 20^3 box (kershaw.re2, Dirichlet pressure on every face) with eps = 0.3,
    rhs = 3 pi^2 sin sin sin at the nodes (kershaw.udf:20-23), x0 = 0, the best of
    `reps` timed solves after a warm-up, as the udf reports.
    BPS5: the pressure defaults (kershaw.par, parReader.cpp:836-841): PGMRES+FLEXIBLE
    restart 15 to 1e-8 relative, p-MG 7-3-1 with RAS + opt. 4th-kind Chebyshev
    degree 3 in FP32; coarse level: one AMG-preconditioned PCG iteration on the
    device (the reference: one BoomerAMG V-cycle on the host).  Run with the FDM in
    FP32 (as the reference) and in BF16 on the matrix cores.
    BP5: PCG, PRECONDITIONER = NONE, tolerance 1e-15 relative, 1000 iterations at
    most (kershaw.udf:47-53)."""
    import numpy as np
    import torch
    import nekrs_amd
    mesh = nekrs_amd.BoxMesh(20, 20, 20, N, "kershaw", 0.3)
    dofs = mesh.E * N ** 3
    f = 3 * np.pi ** 2 * np.sin(np.pi * mesh.x) * np.sin(np.pi * mesh.y) * np.sin(np.pi * mesh.z)
    rhs = torch.from_numpy(f).cuda()

    def timed(el, name):
        best, st_best = None, None
        for rep in range(reps + 1):
            r = rhs.clone()
            x = torch.zeros_like(r)
            torch.cuda.synchronize()
            if rep == reps:
                ranges().push("nrs:kershaw:" + name)
            st = el.solve(r, x)
            if rep == reps:
                ranges().pop()
            if rep > 0 and (best is None or st["seconds"] < best):
                best, st_best = st["seconds"], st
        return best, st_best

    out = {"workload": "Kershaw BPS5/BP5 as in examples/kershaw: 20^3 = %d hex elements N=%d "
                       "eps=0.3, rhs 3 pi^2 sin sin sin, x0 = 0, best of %d solves after a "
                       "warm-up" % (mesh.E, N, reps),
           "dofs": dofs, "published": PUBLISHED_KERSHAW}
    bps5 = dict(precon="multigrid", solver="fgmres", gmresRestart=15, tol=1e-8, max_iter=200,
                levels=[N, 3, 1], smoother="opt4chebyshev+ras", cheby_degree=3, mgPrecision=0,
                coarse="amg", coarse_tol=0.0, coarse_max_iter=1)
    for name, fdm in (("bps5_fdm_f32", 0), ("bps5_fdm_bf16", 1)):
        el = nekrs_amd.Elliptic(mesh, fdmPrecision=fdm, **bps5)
        t, st = timed(el, name)
        out[name] = {"seconds": t, "iterations": st["iterations"], "converged": st["converged"],
                     "precon_seconds": st["precon_seconds"], "dof_per_s": dofs / t,
                     "dof_iter_per_s": dofs * st["iterations"] / t,
                     "speedup_vs_published": {k: v["bps5_s"] / t for k, v in PUBLISHED_KERSHAW.items()}}
        log("Kershaw %s: %d it, %.4f s (A100 published 0.1968 s, 28 it)" %
            (name, st["iterations"], t))
        del el
    el = nekrs_amd.Elliptic(mesh, precon="none", solver="pcg", tol=1e-15, max_iter=1000)
    t, st = timed(el, "bp5")
    out["bp5"] = {"seconds": t, "iterations": st["iterations"],
                  "dof_iter_per_s": dofs * st["iterations"] / t,
                  "speedup_vs_published": {k: dofs * st["iterations"] / t / v["bp5_dof_it_per_s"]
                                           for k, v in PUBLISHED_KERSHAW.items()}}
    log("Kershaw bp5: %d it, %.4f s, %.3g DOF*iter/s (A100 published 3.87e9)" %
        (st["iterations"], t, dofs * st["iterations"] / t))
    return out


def channel_bc(mesh):
    """Face types of a channel on the box: inflow at x = 0 (face 4), outflow at
    x = 1 (face 2), walls at y = 0, 1 (faces 1, 3), symmetry at z = 0, 1 (0, 5)."""
    import numpy as np
    kind = np.array([7, 1, 16, 1, 3, 7], dtype=np.int32)          # by face id
    e2b = np.asarray(mesh.EToB).reshape(-1, 6)
    return np.where(e2b != 0, kind[None, :], 0).astype(np.int32).ravel()


def pressure_rhs_case(mesh, N, steps, warmup, seed=0):
    """tombo::pressureSolve's right-hand side on the operator mesh: one call =
    curl, gs, curl + gradient + combination, gs, weak divergence + mass term,
    surface terms.  Algorithmic bytes (FP64): curl 128 B/node (Ue 24, rx..tz and
    JW 80, W 24), fused rhsU 184 B/node (W 24, invLMM 8, div 8, geometry 80, BF
    24, mue 8, irho 8, rhsU 24), weak divergence 128 B/node (rhsU 24, invLMM 8,
    geometry 80, div 8, rhs 8) -- on trilinear elements each kernel's 80 B/node
    of geometry is 192 B/element of corners instead; each 3-field gather-scatter 52 B per member of
    a shared row (3 x 16 B values + 4 B index) + 4 B per row; the surface
    kernel 16 B per node of each element with a flux face (rhs read and
    written) + 32 B geometry and 24 B field (+8 B invLMM on symmetry faces) per
    node of each flux face."""
    import numpy as np
    import torch
    import nekrs_amd
    t0 = time.perf_counter()
    bc = channel_bc(mesh)
    pr = nekrs_amd.PressureRhs(mesh, bc)
    setup = time.perf_counter() - t0
    L = pr.Nlocal
    rng = np.random.default_rng(seed)
    dv = lambda a: torch.from_numpy(a).cuda()
    Ue, U, BF = (dv(rng.uniform(-1, 1, 3 * L)) for _ in range(3))
    mue, irho = dv(rng.uniform(0.5, 1.5, L)), dv(rng.uniform(0.5, 1.5, L))
    div = dv(rng.uniform(-0.1, 0.1, L))
    rhs = torch.empty(L, dtype=torch.float64, device="cuda")
    g0idt = 1.5 / 1e-3
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(h, rng_name=None):
        for _ in range(warmup):
            h.compute(Ue, U, BF, mue, irho, div, g0idt, True, rhs=rhs)
        torch.cuda.synchronize()
        if rng_name:
            ranges().push(rng_name)
        a.record()
        for _ in range(steps):
            h.compute(Ue, U, BF, mue, irho, div, g0idt, True, rhs=rhs)
        b.record()
        torch.cuda.synchronize()
        if rng_name:
            ranges().pop()
        return a.elapsed_time(b) / steps

    ms = timed(pr, "nrs:pressure_rhs")
    corner = pr.corner_geometry
    info = nekrs_amd.Ogs(mesh.gid).info()
    gs3 = 52.0 * info["members"] + 4.0 * (info["rows"] + 1)
    Np, Nfp = (N + 1) ** 3, (N + 1) ** 2
    e2b = bc.reshape(-1, 6)
    flux = np.isin(e2b, [2, 3, 4, 5, 6, 7])
    sym = np.isin(e2b, [4, 5, 6, 7])
    surf = (16.0 * Np * np.sum(flux.any(1)) + 56.0 * Nfp * flux.sum() + 8.0 * Nfp * sym.sum())
    # each of the three volume kernels reads rx..tz, JW (80 B/node) or, on trilinear
    # elements, the 8 corners (192 B/element) it forms them from
    geo = 3 * (192.0 * mesh.E if corner else 80.0 * L)
    alg = 200.0 * L + geo + 2 * gs3 + surf
    gbs = alg / (ms * 1e-3) / 1e9
    log("pressure rhs (%d elements N=%d, channel, %s geometry): %.4f ms per call, %.2f GB algorithmic, "
        "%.0f GB/s (%.1f %% of HBM peak)" % (mesh.E, N, "corner" if corner else "stored", ms, alg / 1e9,
                                             gbs, 100 * gbs / PEAK_HBM_GBS))
    variants = {k: nekrs_amd.tuned_variant("tombo", N, k, mesh.E) for k in ("curl", "rhs")}
    fused = pr.fused_pairs
    # the same calls without the face fusion, and streaming the stored factors, on the same box
    pr0 = nekrs_amd.PressureRhs(mesh, bc, fusion=False)
    ms0 = timed(pr0)
    del pr0
    prs = nekrs_amd.PressureRhs(mesh, bc, stored_geometry=True)
    ms_stored = timed(prs)
    del prs
    log("pressure rhs: %d face pairs fused; without the fusion %.4f ms per call; streaming the stored "
        "factors %.4f ms per call" % (fused, ms0, ms_stored))
    return {"ms_per_call": ms, "setup_seconds": setup, "algorithmic_bytes": alg,
            "geometry": "corner" if corner else "stored",
            "load_schedule_variants": variants, "fused_face_pairs": fused,
            "unfused_ms_per_call": ms0, "stored_geometry_ms_per_call": ms_stored,
            "achieved_GBps": gbs, "roofline_frac": gbs / PEAK_HBM_GBS, "steps": steps,
            "flux_elements": int(flux.any(1).sum()),
            "workload": "tombo::pressureSolve right-hand side (viscous term on), %d hex elements "
                        "N=%d Kershaw eps=0.3, channel faces (inflow, outflow, walls, symmetry), "
                        "fields U(-1,1) seed %d" % (mesh.E, N, seed)}


# the mixed-precision p-MG of the 48^3 mixed solve, with its FP64 coarse solve (two AMG-preconditioned
# PCG iterations on the assembled order-1 matrix, run on the device so each problem replays as a graph)
BATCH_KW = dict(solver="fcg", precon="multigrid", smoother="opt4chebyshev+asm", cheby_degree=3,
                mgPrecision=0, fdmPrecision=1, coarse="amg", coarse_tol=0.0, coarse_max_iter=2,
                tol=1e-8, max_iter=500)


def box_dims(E):
    """A near-cubic nx*ny*nz box with about E elements (each side >= 2)."""
    nx = max(2, int(round(E ** (1.0 / 3.0))))
    ny = max(2, int(round((E / nx) ** 0.5)))
    nz = max(2, int(round(E / (nx * ny))))
    return nx, ny, nz


def run_batch(probs, iters_per_launch, seed, one_graph=False):
    """Set up one graph batch, solve it (untimed, then timed), then solve the
    same handles one after another (host-driven); returns the measurements.
    one_graph: also time the single-graph mode (all problems as branches of one
    graph; slower on this ROCm, so left out by default)."""
    import numpy as np
    import torch
    import nekrs_amd
    t0 = time.perf_counter()
    b = nekrs_amd.Batch(probs, iters_per_launch=iters_per_launch)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    rng = np.random.default_rng(seed)
    rhs = [torch.from_numpy(rng.uniform(-1, 1, m.Nlocal)).cuda() for m, _ in probs]
    t_mode = {}
    modes = ("one_graph", "per_problem") if one_graph else ("per_problem",)
    xb, st0 = None, None
    for mode in modes:
        t0 = time.perf_counter()
        b.set_mode(mode)
        t_capture = time.perf_counter() - t0
        for rep in range(2):
            r = [t.clone() for t in rhs]
            x = [torch.zeros_like(t) for t in rhs]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = b.solve(r, x)
            torch.cuda.synchronize()
            t_mode[mode] = time.perf_counter() - t0
        info = b.info()
        t_mode[mode + "_launches"] = info["last_launches"]
        t_mode[mode + "_capture_seconds"] = t_capture
        if xb is None:
            xb = [t.cpu().numpy() for t in x]
            st0 = st
        else:   # the two modes give the same iterates
            assert [s_["iterations"] for s_ in st] == [s_["iterations"] for s_ in st0]
            assert all(np.array_equal(t.cpu().numpy(), xb[p]) for p, t in enumerate(x))
    t_batch = min(t_mode[m] for m in modes)
    # the same handles, one problem at a time (host loop per iteration)
    t_seq, seq_it, diff, t_each = 0.0, [], 0.0, []
    for p in range(len(probs)):
        el = b.problem(p)
        r = rhs[p].clone()
        xs = torch.zeros_like(r)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s1 = el.solve(r, xs)
        torch.cuda.synchronize()
        t_each.append(time.perf_counter() - t0)
        t_seq += t_each[-1]
        seq_it.append(s1["iterations"])
        diff = max(diff, float(np.max(np.abs(xs.cpu().numpy() - xb[p]))))
    its = [s["iterations"] for s in st]
    dofs = [m.E * m.N ** 3 for m, _ in probs]
    work = float(sum(d * i for d, i in zip(dofs, its)))
    del b
    return {"problems": len(probs), "setup_seconds": t_setup, "seconds_batched": t_batch,
            "seconds_sequential": t_seq, "speedup_batched_vs_sequential": t_seq / t_batch,
            # the slowest problem solved alone: a floor for any schedule of the batch
            "seconds_slowest_problem_alone": max(t_each),
            "converged": [bool(s["converged"]) for s in st],
            "dof_iterations_per_s_batched": work / t_batch,
            "iterations": its, "iterations_sequential": seq_it,
            "all_converged": all(s["converged"] for s in st),
            "max_abs_diff_batched_vs_sequential": diff,
            "graph_nodes": info["graph_nodes"], "modes": t_mode,
            "iters_per_launch": iters_per_launch, "dofs": dofs}


def block_velocity(m, seed):
    """Three velocity components as ONE block solve (Nfields = 3: vectors
    [3][Nlocal], one gather-scatter of the three fields, inner products summed
    over the fields), flexible CG with the mixed p-MG applied to each field,
    lambda1 = 10, to 1e-8; and the reference's default velocity solver, block
    PCG with Jacobi, on a time step's velocity operator (mu = 0.01, rho bd0/dt =
    1500) with the same right-hand side."""
    import numpy as np
    import torch
    import nekrs_amd
    el = nekrs_amd.Elliptic(m, lambda0=1.0, lambda1=10.0, poisson=False, nfields=3, solver="fcg",
                            precon="multigrid", smoother="opt4chebyshev+asm", cheby_degree=3,
                            mgPrecision=0, fdmPrecision=1, coarse="amg", tol=1e-8, max_iter=500)
    rng = np.random.default_rng(seed)
    rhs = torch.from_numpy(rng.uniform(-1, 1, 3 * el.Nlocal)).cuda()
    for rep in range(2):
        x = torch.zeros_like(rhs)
        torch.cuda.synchronize()
        st = el.solve(rhs.clone(), x)
    dofs = 3 * m.E * m.N ** 3
    out = {"iterations": st["iterations"], "seconds": st["seconds"], "converged": st["converged"],
           "dof_iterations_per_s": dofs * st["iterations"] / st["seconds"]}
    del el
    # the reference's default velocity solver, block PCG with Jacobi (parReader.cpp:996-998,1097),
    # on a time step's velocity Helmholtz operator: mu = 0.01, rho * bd0 / dt = 1.5 / 1e-3
    el = nekrs_amd.Elliptic(m, lambda0=0.01, lambda1=1500.0, poisson=False, nfields=3,
                            solver="pcg", precon="jacobi", tol=1e-8, max_iter=2000)
    for rep in range(2):
        x = torch.zeros_like(rhs)
        torch.cuda.synchronize()
        st = el.solve(rhs.clone(), x)
    out["jacobi_pcg"] = {"operator": "mu=0.01, rho*bd0/dt=1500",
                         "iterations": st["iterations"], "seconds": st["seconds"],
                         "converged": st["converged"],
                         "dof_iterations_per_s": dofs * st["iterations"] / st["seconds"]}
    del el
    return out


def batch_cases(count, seed=0, iters_per_launch=4, block_n=32, one_graph=False):
    """The uneven Helmholtz batch and the 3-component velocity Helmholtz solve."""
    import numpy as np
    import nekrs_amd
    rng = np.random.default_rng(seed)
    probs, desc = [], []
    for p in range(count):
        N = int(rng.choice([3, 5, 7, 9]))
        Et = int(rng.integers(512, 8193))
        lam = float(rng.uniform(0.0, 10.0))
        eps = float(rng.uniform(0.0, 0.3))
        nx, ny, nz = box_dims(Et)
        probs.append((nekrs_amd.BoxMesh(nx, ny, nz, N, "kershaw", eps),
                      dict(BATCH_KW, lambda0=1.0, lambda1=lam, poisson=False)))
        desc.append({"N": N, "E": nx * ny * nz, "box": [nx, ny, nz], "lambda1": lam, "eps": eps})
    log("batch: %d problems, %d elements in all" % (count, sum(d["E"] for d in desc)))
    out = {"uneven": run_batch(probs, iters_per_launch, seed + 1, one_graph)}
    out["uneven"]["cases"] = desc
    u = out["uneven"]
    log("batch: slowest problem alone %.3f s; %d of %d converged within the iteration cap" %
        (u["seconds_slowest_problem_alone"], sum(u["converged"]), len(u["converged"])))
    log("batch: %.3f s batched (one graph %s s, graph per problem %.3f s) vs %.3f s one by one "
        "(x%.2f), %.2f G dof-iterations/s, its %d..%d" %
        (u["seconds_batched"], ("%.3f" % u["modes"]["one_graph"]) if "one_graph" in u["modes"] else "-",
         u["modes"]["per_problem"],
         u["seconds_sequential"], u["speedup_batched_vs_sequential"],
         u["dof_iterations_per_s_batched"] / 1e9, min(u["iterations"]), max(u["iterations"])))
    del probs
    if block_n:
        m = nekrs_amd.BoxMesh(block_n, block_n, block_n, 7, "kershaw", 0.3)
        vel = [(m, dict(BATCH_KW, lambda0=1.0, lambda1=10.0, poisson=False)) for _ in range(3)]
        out["velocity3"] = run_batch(vel, iters_per_launch, seed + 2)
        v = out["velocity3"]
        log("velocity x3 at %d^3 N=7: %.3f s batched vs %.3f s one by one, its %s" %
            (block_n, v["seconds_batched"], v["seconds_sequential"], v["iterations"]))
        out["velocity3_block"] = block_velocity(m, seed + 2)
        vb = out["velocity3_block"]
        log("velocity block (3 fields, one operator, FCG + p-MG per field): %d it, %.3f s; "
            "block Jacobi-PCG (the reference's default) on a time step's velocity operator: "
            "%d it, %.3f s" %
            (vb["iterations"], vb["seconds"], vb["jacobi_pcg"]["iterations"],
             vb["jacobi_pcg"]["seconds"]))
    out["workload"] = ("%d Helmholtz problems (seed %d): N in {3,5,7,9}, E ~ U[512,8192] as "
                       "near-cubic boxes, lambda1 ~ U[0,10], Kershaw eps ~ U[0,0.3]; plus 3 "
                       "velocity components (lambda1=10) at %d^3 N=7 eps 0.3, as a batch and as "
                       "one 3-field block solve; FP32 p-MG, opt. 4th-kind Chebyshev + ASM with "
                       "BF16 FDM, FP64 coarse solve (2 AMG-preconditioned PCG iterations on the "
                       "order-1 matrix; the block solve: p-MG applied per field), FCG to 1e-8 "
                       "relative" %
                       (count, seed, block_n))
    return out


def counter_summary(d, last=0):
    """Mean of each counter per (kernel, grid size) over a rocprofv3 --pmc run
    written with --output-format csv (its *counter_collection.csv files); with
    last > 0 only the last `last` dispatches of each (kernel, grid) count (the
    timed loop, after the setup's tuning launches)."""
    import csv
    import glob
    rows = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                k = (row["Kernel_Name"], int(row["Grid_Size"]), row["Counter_Name"])
                rows.setdefault(k, []).append((int(row["Dispatch_Id"]), float(row["Counter_Value"])))
    acc = {}
    for k, v in rows.items():
        v.sort()
        if last > 0:
            v = v[-last:]
        acc[k] = (sum(x for _, x in v), len(v))
    print("| kernel | grid | counter | dispatches | mean per dispatch |")
    print("|---|---|---|---|---|")
    for (kn, g, cn), (v, n) in sorted(acc.items()):
        short = kn if len(kn) <= 90 else kn[:87] + "..."
        print("| `%s` | %d | %s | %d | %.6g |" % (short, g, cn, n, v / n))


def profile_summary(db, top=40):
    """Markdown summary of a rocprofv3 kernel trace (its rocpd SQLite output):
    per roctx range pushed by this script (nrs:solve:..., nrs:kershaw:...) --
    or over the whole run when there is none -- the kernels grouped by name and
    grid size (so the launches of one kernel on different multigrid levels
    stay apart): calls, total and mean duration."""
    import sqlite3
    con = sqlite3.connect(db)
    kern = list(con.execute("select name, grid_x, workgroup_x, start, end from kernels"))
    regs = []
    try:   # roctx ranges: the message is in the region's extdata ({"message": ...})
        for name, a, b, extra in con.execute("select name, start, end, extdata from regions"):
            try:
                msg = json.loads(extra or "{}").get("message", name)
            except ValueError:
                msg = name
            if str(msg).startswith("nrs:"):
                regs.append((msg, a, b))
    except sqlite3.Error:
        regs = []
    spans = [(n, a, b) for n, a, b in regs] or [("whole run", min(k[3] for k in kern),
                                                 max(k[4] for k in kern))]
    lines = []
    for name, a, b in spans:
        groups = {}
        for kn, gx, wx, s0, s1 in kern:
            if s0 >= a and s1 <= b:
                g = groups.setdefault((kn, gx, wx), [0, 0.0])
                g[0] += 1
                g[1] += (s1 - s0) * 1e-3
        tot = sum(v[1] for v in groups.values())
        lines.append("### %s: %.3f ms of kernels, %.3f ms wall" % (name, tot * 1e-3, (b - a) * 1e-6))
        lines.append("")
        lines.append("| kernel | grid | wg | calls | total ms | mean us | % |")
        lines.append("|---|---|---|---|---|---|---|")
        for (kn, gx, wx), (n, t) in sorted(groups.items(), key=lambda kv: -kv[1][1])[:top]:
            short = kn if len(kn) <= 90 else kn[:87] + "..."
            lines.append("| `%s` | %d | %d | %d | %.3f | %.1f | %.1f |" %
                         (short, gx, wx, n, t * 1e-3, t / n, 100.0 * t / max(tot, 1e-30)))
        lines.append("")
    print("\n".join(lines))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--order", type=int, default=7)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--solve-n", type=int, default=48, help="elements per direction of the solves")
    ap.add_argument("--no-unfused", action="store_true",
                    help="skip the comparison loop of the operator without face fusion")
    ap.add_argument("--operator-fusion", type=int, default=-1,
                    help="operatorFusion of the timed operator (default: the library's, timed at setup)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="set up the timed operator with the autotuner off (face fusion in blocks of "
                         "four, default launch variants): for counter passes, where the profiler's "
                         "per-dispatch serialisation would skew the setup's timings")
    ap.add_argument("--no-solves", action="store_true")
    ap.add_argument("--solve-cases", default="",
                    help="comma-separated names of the p-MG solves to run (default: all)")
    ap.add_argument("--no-pressure-rhs", action="store_true")
    ap.add_argument("--no-plain-poisson", action="store_true",
                    help="skip the 8^3 sine-deformed Jacobi-PCG solve")
    ap.add_argument("--no-kershaw", action="store_true",
                    help="skip the reference's Kershaw BPS5/BP5 benchmarks (20^3, N=7)")
    ap.add_argument("--batch-problems", type=int, default=64)
    ap.add_argument("--batch-iters-per-launch", type=int, default=4)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-one-graph", action="store_true",
                    help="also time the batch's single-graph mode (slower; tested, not timed by default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 5 device copies of 512 MiB (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters), then exit")
    ap.add_argument("--profile-summary", default="",
                    help="print a per-range, per-kernel summary of a rocprofv3 rocpd database "
                         "(.db), or the per-kernel counter means of a --pmc csv output "
                         "directory, and exit")
    ap.add_argument("--last", type=int, default=0,
                    help="with --profile-summary of counters: only the last N dispatches of "
                         "each kernel and grid")
    args = ap.parse_args()
    if args.profile_summary:
        if os.path.isdir(args.profile_summary):
            counter_summary(args.profile_summary, args.last)
        else:
            profile_summary(args.profile_summary)
        return
    if args.copy_calibration:
        import torch
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 5 x %d bytes" % (a.numel() * 8))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the multi-GPU halo exchange of this path is "
              "designed but not built, so --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import nekrs_amd

    nekrs_amd.init(0)
    n, N = args.n, args.order
    t0 = time.perf_counter()
    mesh = nekrs_amd.BoxMesh(n, n, n, N, "kershaw", 0.3)
    okw = {} if args.operator_fusion < 0 else {"operatorFusion": args.operator_fusion}
    if args.no_autotune:
        okw["autotune"] = 0
    el = nekrs_amd.Elliptic(mesh, precon="none", **okw)
    log("setup %.1f s: E=%d N=%d Nlocal=%d Nmasked=%d" %
        (time.perf_counter() - t0, mesh.E, N, el.Nlocal, el.Nmasked))
    # gather-scatter algorithmic bytes: rows with more than one member of the
    # masked handle, 8 B read + 8 B write + 4 B index per member, 4 B offset per row
    mgid = mesh.gid.copy()
    mgid[el.mask_ids.cpu().numpy()] = 0
    info = nekrs_amd.Ogs(mgid).info()
    gs_bytes = 20.0 * info["members"] + 4.0 * (info["rows"] + 1)
    q = torch.from_numpy(np.random.default_rng(0).uniform(-1, 1, el.Nlocal)).cuda()
    Aq = torch.empty_like(q)
    for _ in range(args.warmup):
        el.operator(q, Aq)
    torch.cuda.synchronize()

    el.profile(True)
    torch.cuda.synchronize()
    ranges().push("nrs:operator")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        el.operator(q, Aq)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ranges().pop()
    prof = el.profile_read()
    el.profile(False)
    ms_step = dt / args.steps * 1e3
    dofs = mesh.E * N ** 3
    gdofs = dofs / (ms_step * 1e-3) / 1e9
    ax_ms = prof["ax"][0] / max(prof["ax"][1], 1)
    gs_ms = prof["gs"][0] / max(prof["gs"][1], 1)
    ax_bytes = 64.0 * el.Nlocal
    achieved = ax_bytes / (ax_ms * 1e-3) / 1e9
    log("step %.4f ms (%.2f GDOF/s); Ax %.4f ms (%.0f GB/s), gs %.4f ms" %
        (ms_step, gdofs, ax_ms, achieved, gs_ms))

    # the device copy bandwidth on this box (512 MiB read + 512 MiB written per copy):
    # the practical ceiling a streaming kernel can reach, reported beside the nominal peak
    ca = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
    cb = torch.empty_like(ca)
    for _ in range(3):
        cb.copy_(ca)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(20):
        cb.copy_(ca)
    ev1.record()
    torch.cuda.synchronize()
    copy_gbs = 2 * ca.numel() * 8 / (ev0.elapsed_time(ev1) / 20 * 1e-3) / 1e9
    del ca, cb
    log("device copy: %.0f GB/s" % copy_gbs)

    Nlocal = el.Nlocal
    fused_pairs = el.fused_pairs
    del el
    fusion = {"fused_pairs": fused_pairs, "fused_ms_per_step": ms_step}
    if not args.no_unfused:
        # the same loop without the operator face fusion, on the same box, for comparison
        el0 = nekrs_amd.Elliptic(mesh, precon="none", operatorFusion=0)
        for _ in range(args.warmup):
            el0.operator(q, Aq)
        el0.profile(True)
        torch.cuda.synchronize()
        ranges().push("nrs:operator_unfused")
        t0 = time.perf_counter()
        for _ in range(args.steps):
            el0.operator(q, Aq)
        torch.cuda.synchronize()
        ms0 = (time.perf_counter() - t0) / args.steps * 1e3
        ranges().pop()
        p0 = el0.profile_read()
        el0.profile(False)
        del el0
        fusion.update({"unfused_ms_per_step": ms0, "unfused_ax_ms": p0["ax"][0] / max(p0["ax"][1], 1),
                       "unfused_gs_ms": p0["gs"][0] / max(p0["gs"][1], 1)})
        log("without the face fusion: step %.4f ms (%.2f GDOF/s); Ax %.4f ms, gs %.4f ms" %
            (ms0, dofs / (ms0 * 1e-3) / 1e9, fusion["unfused_ax_ms"], fusion["unfused_gs_ms"]))
    prhs = None
    if not args.no_pressure_rhs:
        prhs = pressure_rhs_case(mesh, N, 20, 3)
    plain = None
    if not args.no_plain_poisson:
        plain = plain_poisson_case(cpu=not args.no_cpu_baseline)
    kershaw = None
    if not args.no_kershaw:
        log("Kershaw BPS5 / BP5 at 20^3 ...")
        kershaw = kershaw_cases(N)
    solves = None
    if not args.no_solves:
        log("p-MG solves at %d^3 ..." % args.solve_n)
        solves = solve_cases(args.solve_n, N, only=[c for c in args.solve_cases.split(",") if c])

    batch = None
    if not args.no_batch:
        log("batched solves ...")
        batch = batch_cases(args.batch_problems, 0, args.batch_iters_per_launch, n,
                            one_graph=args.batch_one_graph)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(n, N, args.cpu_seconds, 0)
        log("cpu baseline %.3f GDOF/s on %d cores" % (cpu["value"], cpu["cores"]))

    out = {
        "metric": METRIC,
        "value": gdofs,
        "unit": "GDOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "Ax+gs FP64, %d^3=%d hex elements N=%d Kershaw eps=0.3, "
                               "q~U(-1,1) seed 0" % (n, mesh.E, N),
                   "elements": mesh.E, "order": N, "local_nodes": Nlocal,
                   "kernel_ms": {"ax": ax_ms, "gs": gs_ms},
                   "gs_rows": info["rows"], "gs_members": info["members"],
                   "gs_achieved_GBps": gs_bytes / (gs_ms * 1e-3) / 1e9,
                   "axgs_algorithmic_GBps": (ax_bytes + gs_bytes) / (ms_step * 1e-3) / 1e9,
                   "axgs_roofline_frac": (ax_bytes + gs_bytes) / (ms_step * 1e-3) / 1e9 /
                   PEAK_HBM_GBS,
                   "operator_face_fusion": fusion,
                   "device_copy_GBps": copy_gbs,
                   "ax_frac_of_device_copy": achieved / copy_gbs},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic_from_profiles()},
        "cpu_baseline": cpu,
        "pressure_rhs": prhs,
        "plain_poisson": plain,
        "solves": solves,
        "kershaw": kershaw,
        "batch": batch,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
