"""Oracle for the p-multigrid preconditioned (flexible) PCG.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  A numpy driver over the
oracle's C kernels that restates the reference's multigrid:

  V-cycle                   src/solvers/elliptic/MG/MGSolver.cpp:167-193
  smooth / residual         MG/ellipticMultiGridLevel.cpp:37-41, 85-277
  coarsen / prolongate      MG/ellipticMultiGridLevel.cpp:43-83
  Chebyshev bounds          MG/ellipticMultiGridLevelSetup.cpp:109-180, 292-453
  level meshes              src/mesh/meshSetup.cpp:293-376 (coordinates
                            interpolated from the fine mesh, geometric factors
                            recomputed, FP32 copies)
  preconditioner casts      src/solvers/elliptic/ellipticPreconditioner.cpp:57-62
  PCG / flexible PCG        src/solvers/elliptic/linearSolver/PCG.cpp:85-203

MULTIGRID SCHEDULE strings: parse_mg_schedule restates parseMultigridSchedule
(MG/parseMultigridSchedule.cpp:7-131), the levels determineMGLevels takes from
it (MG/determineMGLevels.cpp:22-40) and the per-level degrees the level setup
reads (MG/ellipticMultiGridLevelSetup.cpp:160-167).

Smoothers: Chebyshev of the 1st kind (ellipticMultiGridLevel.cpp:144-210) or
the optimal 4th kind (:212-270, betas of MG/optimalCoeffs.cpp) around S =
Jacobi (1/diag A) or the overlapping Schwarz / FDM smoother (oracle/fdm.py,
ASM or RAS); pure ASM / RAS smoothing (:85-103: no post-smoothing when x is
nonzero).

Level arithmetic is FP32 (pfloat) or FP64; the coarsest level (N=1) is solved
by Jacobi-PCG in FP64 to a relative tolerance, where the reference calls
BoomerAMG.  lambda_max(S A) comes from a 10-step Arnoldi started from a
splitmix64 vector (the reference seeds from std::random_device, which is not
reproducible) and numpy's eigvals (LAPACK geev, as the reference's dgeev).
"""
import numpy as np

from . import (BoxMesh, Elliptic, GWJ, Ogs, gll, interp, lib, mask_ids, masked_gids,
               transfer_R, build_diag)
from . import fdm as _fdm

# optimal 4th-kind Chebyshev betas, degrees 1..8 (MG/optimalCoeffs.cpp:4-215;
# Lottes, arXiv:2202.08830)
OPT4_BETAS = {
    1: [1.12500000000000],
    2: [1.02387287570313, 1.26408905371085],
    3: [1.00842544782028, 1.08867839208730, 1.33753125909618],
    4: [1.00391310427285, 1.04035811188593, 1.14863498546254, 1.38268869241000],
    5: [1.00212930146164, 1.02173711549260, 1.07872433192603, 1.19810065292663, 1.41322542791682],
    6: [1.00128517255940, 1.01304293035233, 1.04678215124113, 1.11616489419675, 1.23829020218444, 1.43524297106744],
    7: [1.00083464397912, 1.00843949430122, 1.03008707768713, 1.07408384092003, 1.15036186707366, 1.27116474046139, 1.45186658649364],
    8: [1.00057246631197, 1.00577427662415, 1.02050187922941, 1.05019803444565, 1.10115572984941, 1.18086042806856, 1.29838585382576, 1.46486073151099]}

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _pieces(s, dlim):
    """serializeString (src/core/nrssys.hpp:148-160): white space removed,
    split at dlim, empty pieces dropped."""
    s = "".join(c for c in s if not c.isspace())
    return [p for p in s.split(dlim) if p]


def _stoi(token):
    """std::stoi of the second '='-piece: optional sign, leading digits."""
    p = _pieces(token, "=")
    if len(p) < 2:
        raise ValueError("no value in token '%s'" % token)
    v, i = p[1], 0
    if i < len(v) and v[i] in "+-":
        i += 1
    j = i
    while j < len(v) and v[j].isdigit():
        j += 1
    if j == i:
        raise ValueError("no value in token '%s'" % token)
    return int(v[:j])


def parse_mg_schedule(schedule, N, default_degree):
    """(levels, pre, post, errors): levels = distinct down-leg orders, finest
    first; pre[i] / post[i] = Chebyshev degree before restriction / after
    prolongation on level i (0 on the coarsest); errors = the reference's
    messages (empty when valid).  The degree before restriction is the one on
    the schedule's up leg (DownLegChebyshevDegree = scheduleMap[{order, false}],
    ellipticMultiGridLevelSetup.cpp:165-166), after prolongation the down leg's."""
    INVALID = -(2 ** 31)
    err = []
    entries = _pieces(schedule, ",")
    min_order = 2 ** 31 - 1
    for e in entries:
        for t in _pieces(e, "+"):
            if "p" in t:
                min_order = min(min_order, _stoi(t))
    sched = {}
    down_orders = []
    down, prev = True, 2 ** 31 - 1
    for e in entries:
        order, degree = INVALID, INVALID
        for t in _pieces(e, "+"):
            if "p" in t:                      # as there: "sweeps" contains a 'p'
                order = _stoi(t)
                if order > prev:
                    down = False
                if down:
                    down_orders.append(order)
                prev = order
            elif "degree" in t:
                degree = _stoi(t)
            else:
                err.append("unknown token '%s' in schedule '%s'!" % (t, schedule))
        if order == INVALID:
            err.append("order not specified in " + e)
            continue
        if degree == INVALID and order != min_order:
            degree = default_degree
        sched[(order, down)] = degree
    dset = {o for (o, d) in sched if d and o > min_order}
    uset = {o for (o, d) in sched if not d and o > min_order}
    if dset != uset:
        err.append("down leg and up leg orders must be identical!")
    for (o, d), g in sched.items():
        if o != min_order and g == INVALID:
            err.append("degree not specified for order %d!" % o)
    for l in down_orders:
        if l < 1:
            err.append("encountered level = %d < 1!" % l)
    if not down_orders:
        err.append("no levels")
    elif down_orders[0] != N:
        err.append("finest degree %d does not match polynomial degree %d!" % (down_orders[0], N))
    for (o, d), g in sched.items():
        if o != min_order and g != INVALID and not 1 <= g <= 16:
            err.append("degree %d of order %d outside 1..16!" % (g, o))
    levels = sorted({o for (o, d) in sched if d}, reverse=True)
    if len(levels) > 8:
        err.append("more than 8 levels")
    if err:
        return levels, [], [], err
    pre = [sched[(o, False)] for o in levels[:-1]] + [0]
    post = [sched[(o, True)] for o in levels[:-1]] + [0]
    return levels, pre, post, []


def splitmix_uniform(seed, n):
    """n draws of splitmix64 (state advanced by the golden-ratio increment)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def interp_coords(Nf, E, xf, Nc):
    zf, _ = gll(Nf)
    zc, _ = gll(Nc)
    I = interp(zf, zc)
    u = xf.reshape(E, Nf + 1, Nf + 1, Nf + 1)
    return np.einsum("ck,bj,ai,ekji->ecba", I, I, I, u, optimize=True).ravel()


class Level:
    pass


class MG:
    """p-multigrid V-cycle on the levels `orders` of a box mesh."""

    def __init__(self, mesh, orders, T=np.float32, degree=2, min_factor=0.1, max_factor=1.1,
                 coarse_tol=1e-2, coarse_maxit=200, seed=0, cheby="cheby1", S="jacobi",
                 fmt=None, lambda0=1.0, lambda1=0.0, poisson=True, coarse="pcg",
                 pre_degree=None, post_degree=None):
        """cheby: "cheby1", "opt4" or None (pure Schwarz); S: "jacobi", "asm", "ras";
        fmt: None, "bf16" or "fp16" for the FDM contractions; lambda0/lambda1/poisson:
        the (constant-coefficient) operator; coarse: "pcg" (FP64 Jacobi-PCG) or
        "smooth" (MULTIGRID COARSE SOLVE = FALSE: the coarsest level is smoothed,
        degree 3, ellipticMultiGridSetup.cpp:315-320) or "smooth_jacobi" (smoothed
        with Chebyshev-Jacobi whatever S is; an extension); pre_degree /
        post_degree: per-level Chebyshev degrees before restriction / after
        prolongation (0 or None: `degree`), as a MULTIGRID SCHEDULE sets them."""
        assert orders[0] == mesh.N
        assert cheby in ("cheby1", "opt4", None) and S in ("jacobi", "asm", "ras")
        assert cheby is not None or S != "jacobi"
        assert coarse in ("pcg", "smooth", "smooth_jacobi", "amg")
        self.T = T
        self.cheby, self.S = cheby, S
        self.lambda0, self.lambda1, self.poisson = lambda0, lambda1, poisson
        self.coarse = coarse
        self.coarse_tol, self.coarse_maxit = coarse_tol, coarse_maxit
        self.levels = []
        # all-Neumann (ellipticSetup.cpp:187-206): the coarse problem is singular;
        # its right-hand side is made consistent and the AMG's dense level is the
        # pseudo-inverse (the reference passes the flag to BoomerAMG,
        # coarseLevel.cpp:113-146)
        eb = np.asarray(mesh.EToB)
        self.null_space = bool(poisson) and not np.any((eb > 0) & (eb != 2))
        x, y, z = mesh.x, mesh.y, mesh.z
        Nprev = mesh.N
        for li, N in enumerate(orders):
            L = Level()
            L.N, L.Np, L.E = N, (N + 1) ** 3, mesh.E
            L.isCoarse = li == len(orders) - 1
            if li > 0:
                x = interp_coords(Nprev, mesh.E, x, N)
                y = interp_coords(Nprev, mesh.E, y, N)
                z = interp_coords(Nprev, mesh.E, z, N)
            # structured ids and BCs at order N (the fine mesh's boundary types and periodicity)
            m = BoxMesh(mesh.nx, mesh.ny, mesh.nz, N, bc=getattr(mesh, "bc", "dirichlet"),
                        periodic=getattr(mesh, "periodic", ""))
            m.x, m.y, m.z = x, y, z
            _, wq = gll(N)
            m.ggeo = np.zeros(m.E * 7 * m.Np)
            lib().orc_geofacs(N, m.E, m.D.ravel().copy(), wq, x, y, z, m.ggeo, None)
            L.mesh = m
            L.el = Elliptic(m, lambda0, lambda1, poisson, zero_mean=False)   # masked ogs, mask, invDegree
            L.ggeoT = m.ggeo.astype(T)
            L.DT = m.D.astype(T)
            L.invDegreeT = L.el.invDegree.astype(T)
            if li > 0:
                L.R = transfer_R(Nprev, N).astype(T)
            if li == 0 and S != "jacobi":
                fine = (N, x, y, z, wq, L.el.ogs_full)
            if not L.isCoarse or coarse in ("smooth", "smooth_jacobi"):
                if S == "jacobi" or (L.isCoarse and coarse == "smooth_jacobi"):
                    d = build_diag(N, m.ggeo, m.D, lambda0, lambda1, poisson)
                    d = L.el.ogs.gs(d, "add")
                    L.invDiagT = (1.0 / d).astype(T)
                else:
                    L.schwarz = _fdm.Schwarz(m, fine, L.el.maskIds, L.el.ogs, L.el.invDegree,
                                             mode=S, T=T, fmt=fmt, poisson=poisson,
                                             lambda0=lambda0, lambda1=lambda1)
                L.degree = 3 if L.isCoarse else degree
                L.deg_pre = L.deg_post = L.degree
                if not L.isCoarse:
                    if pre_degree is not None and pre_degree[li]:
                        L.deg_pre = pre_degree[li]
                    if post_degree is not None and post_degree[li]:
                        L.deg_post = post_degree[li]
            self.levels.append(L)
            Nprev = N
        for L in self.levels:
            if (not L.isCoarse or coarse in ("smooth", "smooth_jacobi")) and cheby is not None:
                L.rho = self._max_eig(L, seed)
                L.lmax, L.lmin = max_factor * L.rho, min_factor * L.rho
        cl = self.levels[-1]
        cl.invDiag64 = cl.el.inv_diag()
        if coarse == "amg":
            from . import amg as _amg
            g = cl.mesh.gid.copy()
            g[cl.el.maskIds] = 0
            cl.amg_rows = _amg.row_numbers(g)
            n = int(cl.amg_rows.max()) + 1
            cl.amg_first = np.zeros(n, dtype=np.int64)
            seen = np.zeros(n, dtype=bool)
            for i, r in enumerate(cl.amg_rows.tolist()):
                if r >= 0 and not seen[r]:
                    seen[r] = True
                    cl.amg_first[r] = i
            cl.amg_A = _amg.assemble(cl, lambda u: self.op_local64(cl, u), cl.amg_rows)
            cl.amg = _amg.AMG(cl.amg_A, null_space=self.null_space)

    # ---- level operators
    def op_local64(self, L, q):
        """FP64 element operator of the level (no mask, no gather-scatter)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        Aq = np.zeros_like(q)
        l0 = np.full(1, self.lambda0)
        l1 = np.full(1, self.lambda1)
        lib().orc_ax_f64(L.N, L.E, None, L.mesh.ggeo, L.mesh.D.ravel().copy(), l0, l1,
                         int(self.poisson), 0, q, Aq)
        return Aq

    def op(self, L, q):
        q = np.ascontiguousarray(q, dtype=self.T)
        fn = lib().orc_ax_f64 if self.T == np.float64 else lib().orc_ax_f32
        Aq = np.zeros_like(q)
        l0 = np.full(1, self.lambda0, dtype=self.T)
        l1 = np.full(1, self.lambda1, dtype=self.T)
        fn(L.N, L.E, None, L.ggeoT, L.DT.ravel().copy(), l0, l1, int(self.poisson), 0, q, Aq)
        Aq[L.el.maskIds] = 0
        return L.el.ogs.gs(Aq, "add")

    def _max_eig(self, L, seed):
        k = 10
        n = L.E * L.Np
        vx = splitmix_uniform(seed + 0x5EED * (L.N + 1), n)
        vx = L.el.ogs_full.gs(vx, "add")
        vx[L.el.maskIds] = 0.0
        w = L.el.invDegree
        nv = np.sqrt(np.sum(w * vx * vx))
        if not nv > 0.0:
            return 1.0              # no unknowns on this level (the reference divides by 0)
        V = [vx / nv]
        H = np.zeros((k, k))
        m = k
        for j in range(k):
            av = self.op(L, V[j].astype(self.T))
            v = self.smoother_S(L, av).astype(np.float64)
            hn = 0.0
            for i in range(j + 1):
                hij = np.sum(w * V[i] * v)
                v = v - hij * V[i]
                H[i, j] = hij
                hn += hij * hij
            if j + 1 < k:
                nj = np.sqrt(np.sum(w * v * v))
                # breakdown: invariant Krylov space, its leading block is exact
                # (the reference normalises the rounding residue instead)
                if nj <= 1e-8 * np.sqrt(hn + nj * nj):
                    m = j + 1
                    break
                v = v / nj
                H[j + 1, j] = nj
            V.append(v)
        return float(np.max(np.abs(np.linalg.eigvals(H[:m, :m]))))

    def smoother_S(self, L, v):
        """pMGLevel::smoother (ellipticMultiGridLevel.cpp:104-111)."""
        if getattr(L, "schwarz", None) is None:
            return (L.invDiagT * v).astype(self.T)
        return L.schwarz.apply(v)

    def smooth(self, L, rhs, x, x_is_zero):
        """pMGLevel::smooth (ellipticMultiGridLevel.cpp:85-103)."""
        if self.cheby is None:
            if x_is_zero:
                x[:] = L.schwarz.apply(rhs)
            return x
        if self.cheby == "opt4":
            return self.smooth_opt4(L, rhs, x, x_is_zero)
        return self.smooth_cheby1(L, rhs, x, x_is_zero)

    # smoothChebyshev (ellipticMultiGridLevel.cpp:144-210); scalars are formed
    # in FP64 and stored as pfloat, as there
    def smooth_cheby1(self, L, rhs, x, x_is_zero):
        T = self.T
        deg = L.deg_pre if x_is_zero else L.deg_post
        theta = T(0.5 * (L.lmax + L.lmin))
        delta = T(0.5 * (L.lmax - L.lmin))
        invTheta = T(1.0 / float(theta))
        sigma = T(theta / delta)
        rho_n = T(1.0 / float(sigma))
        if x_is_zero:
            x[:] = 0
            res = rhs.copy()
        else:
            res = (rhs - self.op(L, x)).astype(T)
        res = self.smoother_S(L, res)
        d = (invTheta * res).astype(T)
        for _ in range(1, deg):
            Ad = self.smoother_S(L, self.op(L, d))
            rhoSave = rho_n
            rho_n = T(1.0 / (2.0 * float(sigma) - float(rho_n)))
            rC = T(2.0 * float(rho_n) / float(delta))
            dC = T(rho_n * rhoSave)
            x += d
            res = (res - Ad).astype(T)
            d = (dC * d + rC * res).astype(T)
        x += d
        x[L.el.maskIds] = 0
        return x

    # smoothFourthKindChebyshev (ellipticMultiGridLevel.cpp:212-270), rho = lambda1
    def smooth_opt4(self, L, rhs, x, x_is_zero):
        T = self.T
        deg = L.deg_pre if x_is_zero else L.deg_post
        betas = [T(b) for b in OPT4_BETAS[deg]]
        rho = L.lmax
        if x_is_zero:
            x[:] = 0
            res = rhs.copy()
        else:
            res = (rhs - self.op(L, x)).astype(T)
        Ad = self.smoother_S(L, res)
        d = (T(4.0 / (3.0 * rho)) * Ad).astype(T)
        for k in range(1, deg):
            Ad = self.op(L, d)
            x += (betas[k - 1] * d).astype(T)
            res = (res - Ad).astype(T)
            Ad = self.smoother_S(L, res)
            dC = T((2.0 * k - 1.0) / (2.0 * k + 3.0))
            rC = T((8.0 * k + 4.0) / ((2.0 * k + 3.0) * rho))
            d = (rC * Ad + dC * d).astype(T)
        x += (betas[-1] * d).astype(T)
        x[L.el.maskIds] = 0
        return x

    def coarse_solve_amg(self, L, rhs):
        """PCG on the assembled order-1 matrix, preconditioned by one AMG
        V-cycle, from 0 to coarse_tol relative (the library's coarse AMG
        mode; scalars as in PCG.cpp:85-203)."""
        A, H = L.amg_A, L.amg
        b = rhs.astype(np.float64)[L.amg_first]
        if self.null_space:
            b = b - np.sum(b) / b.size
        x = np.zeros_like(b)
        r = b.copy()
        r0 = np.sqrt(r @ r)
        if not r0 > 0:
            return np.zeros(L.E * L.Np, dtype=self.T), 0
        tol = self.coarse_tol * r0
        p = np.zeros_like(b)
        rdotz_old, it = 0.0, 0
        while it < self.coarse_maxit:
            z = H.vcycle(r)
            rdotz = r @ z
            beta = rdotz / rdotz_old if it > 0 else 0.0
            p = z + beta * p
            Ap = A @ p
            alpha = rdotz / (p @ Ap + 1e-300)
            r = r - alpha * Ap
            x = x + alpha * p
            rdotz_old = rdotz
            it += 1
            if np.sqrt(r @ r) <= tol:
                break
        xe = np.where(L.amg_rows >= 0, x[np.maximum(L.amg_rows, 0)], 0.0)
        return xe.astype(self.T), it

    def coarse_solve(self, L, rhs):
        if self.coarse == "amg":
            return self.coarse_solve_amg(L, rhs)
        el = L.el
        r = rhs.astype(np.float64).copy()
        x = np.zeros_like(r)
        w = el.invDegree
        if self.null_space:         # mean over the unique nodes removed
            r = r - np.sum(w * r) / np.sum(w)
        r0 = np.sqrt(np.sum(w * r * r) / el.volume)
        if not r0 > 0:
            return x.astype(self.T), 0
        import ctypes as C
        rd = C.c_double(r0)
        it = lib().orc_pcg(C.byref(el._c), 1, L.invDiag64.ctypes.data, 0, r, x,
                           self.coarse_tol * r0, self.coarse_maxit, C.byref(rd), None)
        return x.astype(self.T), it

    def vcycle(self, k, rhs):
        L = self.levels[k]
        if L.isCoarse:
            if self.coarse in ("smooth", "smooth_jacobi"):
                x = np.zeros(L.E * L.Np, dtype=self.T)
                return self.smooth(L, rhs, x, True)
            x, self.last_coarse_it = self.coarse_solve(L, rhs)
            return x
        C_ = self.levels[k + 1]
        x = np.zeros(L.E * L.Np, dtype=self.T)
        self.smooth(L, rhs, x, True)
        res = (rhs - self.op(L, x)).astype(self.T)
        res = (res * L.invDegreeT).astype(self.T)
        rc = self._coarsen(L.N, C_.N, C_.R, res, L.E)
        rc = C_.el.ogs.gs(rc, "add")
        rc[C_.el.maskIds] = 0
        xc = self.vcycle(k + 1, rc)
        x = (x + self._prolong(L.N, C_.N, C_.R, xc, L.E)).astype(self.T)
        self.smooth(L, rhs, x, False)
        return x

    def _coarsen(self, Nf, Nc, R, qf, E):
        u = qf.reshape(E, Nf + 1, Nf + 1, Nf + 1)
        return np.einsum("ck,bj,ai,ekji->ecba", R, R, R, u, optimize=True).astype(self.T).ravel()

    def _prolong(self, Nf, Nc, R, qc, E):
        u = qc.reshape(E, Nc + 1, Nc + 1, Nc + 1)
        return np.einsum("ck,bj,ai,ekji->ecba", R.T, R.T, R.T, u, optimize=True).astype(self.T).ravel()

    def apply(self, r):
        """z = V-cycle(r) with FP64 in/out (ellipticPreconditioner.cpp:57-62)."""
        return self.vcycle(0, r.astype(self.T)).astype(np.float64)


def fcg_solve(el, rhs, precon, flexible=True, tol=1e-8, maxit=500, x0=None):
    """ellipticSolve + pcg with a preconditioner callback (FP64 numpy)."""
    m = el.mesh
    x = np.zeros(m.Nlocal) if x0 is None else x0.copy()
    from . import ax
    Ax0 = ax(m.N, m.ggeo, m.D, x, el.lambda0, el.lambda1, el.poisson)
    r = rhs - Ax0
    if el.zero_mean:                                   # ellipticSolve.cpp:101
        r = zero_mean(r)
    r[el.maskIds] = 0
    r = el.ogs.gs(r, "add")
    xs, st = pcg(el, r, precon, flexible, tol, maxit)
    xs = xs + x
    if el.zero_mean:                                   # ellipticSolve.cpp:188-189
        xs = zero_mean(xs)
    return xs, st


def zero_mean(q):
    """ellipticZeroMean (ellipticZeroMean.cpp:31-45): q - sum(q) / size, the sum
    over every local node."""
    return q - np.sum(q) / q.size


def pcg(el, r, precon, flexible=True, tol=1e-8, maxit=500):
    """pcg (PCG.cpp:85-203) from x = 0 on the assembled, masked residual r;
    tol relative to its 1/volume-scaled weighted norm."""
    w = el.invDegree
    vol = el.volume
    res0 = np.sqrt(np.sum(w * r * r) / vol)
    tol_abs = tol * res0
    p = np.zeros_like(r)
    Ap = np.zeros_like(r)
    rdotz1 = 0.0
    alpha = 0.0
    it = 0
    hist = []
    xs = np.zeros_like(r)
    while True:
        it += 1
        z = precon(r)
        if el.zero_mean:                               # ellipticPreconditioner.cpp:82-83
            z = zero_mean(z)
        rdotz2 = rdotz1
        rdotz1 = np.sum(w * r * z)
        beta = 0.0
        if it > 1:
            beta = rdotz1 / rdotz2
            if flexible:
                beta = -alpha * np.sum(w * z * Ap) / rdotz2
        p = z + beta * p
        Ap = el.operator(p)
        alpha = rdotz1 / (np.sum(w * p * Ap) + 1e-300)
        r = r - alpha * Ap
        xs = xs + alpha * p
        rdotr = np.sqrt(np.sum(w * r * r) / vol)
        hist.append(rdotr)
        if not (rdotr > tol_abs and it < maxit):
            break
    return xs, {"iterations": it, "res0": res0, "res": rdotr, "history": np.array(hist)}


def gmres_solve(el, rhs, precon, flexible=True, restart=15, tol=1e-8, maxit=500, nfields=1):
    """ellipticSolve + pgmres (PGMRES.cpp:111-340) with numpy: classical
    Gram-Schmidt with invDegree-weighted inner products, Givens rotations,
    restart; tol relative to the 1/volume-scaled initial norm.  nfields > 1:
    the same operator on each field, inner products summed over the fields."""
    L = el.mesh.Nlocal
    mask = np.isin(np.arange(L), el.maskIds)
    if el.zero_mean:
        rhs = zero_mean(np.asarray(rhs, dtype=np.float64))
    r = np.concatenate([el.ogs.gs(np.where(mask, 0.0, rhs[f * L:(f + 1) * L]), "add")
                        for f in range(nfields)])
    x, st = gmres(el, r, precon, flexible, restart, tol, maxit, nfields)
    if el.zero_mean:
        x = zero_mean(x)
    return x, st


def gmres(el, r, precon, flexible=True, restart=15, tol=1e-8, maxit=500, nfields=1):
    """pgmres from x = 0 on the assembled, masked residual r (all fields)."""
    L = el.mesh.Nlocal
    A = lambda v: np.concatenate([el.operator(v[f * L:(f + 1) * L]) for f in range(nfields)])
    zm = zero_mean if el.zero_mean else (lambda q: q)
    M = lambda v: np.concatenate([zm(precon(v[f * L:(f + 1) * L])) for f in range(nfields)])
    w = np.tile(el.invDegree, nfields)
    rnf = 1.0 / el.volume
    dot = lambda a, b: float(np.sum(w * a * b))
    res0 = np.sqrt(dot(r, r) * rnf)
    tol_abs = tol * res0
    x = np.zeros_like(r)
    b = r.copy()
    nr = res0 / np.sqrt(rnf)
    err = res0
    it = 0
    if not nr > 0:      # zero residual: x = 0 (the library's early return; the reference divides by 0)
        return x, {"iterations": 0, "res0": res0, "res": res0}

    def update(k, H, s, V, Z):
        y = np.zeros(k)
        for a in range(k - 1, -1, -1):
            y[a] = s[a]
            for c in range(a + 1, k):
                y[a] -= H[a, c] * y[c]
            y[a] /= H[a, a]
        if flexible:
            return sum(y[j] * Z[j] for j in range(k))
        return M(sum(y[j] * V[j] for j in range(k)))

    while it < maxit:
        H = np.zeros((restart + 1, restart + 1))
        cs, sn = np.zeros(restart), np.zeros(restart)
        s = np.zeros(restart + 1)
        s[0] = nr
        V = [r / nr]
        Z = []
        done = False
        for i in range(restart):
            z = M(V[i])
            Z.append(z)
            wv = A(z)
            y = np.array([dot(V[j], wv) for j in range(i + 1)])
            for j in range(i + 1):
                wv = wv - y[j] * V[j]
            nw = np.sqrt(dot(wv, wv))
            H[i + 1, i] = nw
            if i < restart - 1:
                V.append(wv / nw)
            H[:i + 1, i] = y
            for k in range(i):
                h1, h2 = H[k, i], H[k + 1, i]
                H[k, i] = cs[k] * h1 + sn[k] * h2
                H[k + 1, i] = -sn[k] * h1 + cs[k] * h2
            h1, h2 = H[i, i], H[i + 1, i]
            hr = np.sqrt(h1 * h1 + h2 * h2)
            cs[i], sn[i] = h1 / hr, h2 / hr
            H[i, i] = cs[i] * h1 + sn[i] * h2
            H[i + 1, i] = 0.0
            s[i + 1] = -sn[i] * s[i]
            s[i] = cs[i] * s[i]
            it += 1
            err = abs(s[i + 1]) * np.sqrt(rnf)
            if err < tol_abs or it == maxit:
                x = x + update(i + 1, H, s, V, Z)
                done = True
                break
        if done:
            break
        x = x + update(restart, H, s, V, Z)
        r = b - A(x)
        nr = np.sqrt(dot(r, r))
        err = nr * np.sqrt(rnf)
        if err <= tol_abs:
            break
    return x, {"iterations": it, "res0": res0, "res": err}

