"""Oracle for the overlapping Schwarz smoother (ASM / RAS) with fast
diagonalisation (FDM) on the extended Nq_e = Nq + 2 element boxes.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  numpy restatement of

  element lengths        src/solvers/elliptic/MG/ellipticMultiGridSchwarz.cpp:58-276
                         (harmonic_mean_element_length, compute_element_lengths)
  boundary flags         :278-298 (compute_element_boundary_conditions)
  1D stiffness / mass    :308-452 (compute_1d_stiffness_matrix, compute_1d_mass_matrix)
  eigenproblem           :455-556 (solve_generalized_ev: LAPACK ?sygv, itype 1,
                         upper triangle; row zeroing for Dirichlet/Neumann ends)
  operators              :558-641 (gen_operators: S per direction, invL)
  extended-mesh mask     :643-743 (create_extended_mesh: edge nodes and
                         Dirichlet faces of the order N+2 mesh are masked)
  Schwarz weights        :798-835 (generate_weights, extrude, to_reg)
  smoother application   :1056-1156 (smoothSchwarz) with kernels/elliptic/
                         preFDM.c:3-98, fusedFDM.c:3-224, postFDM.c:3-155

following the reference's data flow step by step: fill the extended box, gather-
scatter it over the extended mesh, take the neighbour's overlap layer back out,
apply (Sz x Sy x Sx) invL (Sz x Sy x Sx)^T, and either restrict (RAS) or
exchange the overlap results (ASM) before the regular gather-scatter.

The generalized eigenproblem is solved by scipy.linalg.eigh(driver="gv"), the
same LAPACK routine (?sygv) the reference calls; eigenvector signs may differ
between LAPACK builds, which leaves S invL S^T unchanged.

Precision: the setup is FP64, the operators are stored in the smoother type T
(FP32 as pfloat, or FP64).  ``fmt`` = "bf16" / "fp16" emulates FDM contractions
whose inputs (S and each contracted work array) are rounded to that format with
FP32 accumulation -- the mixed-precision variant the GPU path offers.
"""
import os

import numpy as np

from . import BoxMesh, Ogs, DIRICHLET, gll

TOL_LEN = 1e-12
EPS_DIAG = 1e-5
BIG = 1000000000


# ------------------------------------------------------------ rounding
def round_fmt(x, fmt, operator=False):
    """Round float32 values to the 16-bit format (nearest even) and back.

    bf16: 8-bit significand.  fp16: the operators (operator=True) are rounded
    to IEEE half; data are rounded to an 11-bit significand with an unbounded
    exponent, which is what half precision gives for data scaled by a power of
    two per element and contraction (block floating point) as the GPU path does."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if fmt is None or fmt == "f32":
        return x
    if fmt == "fp16":
        if operator:
            return x.astype(np.float16).astype(np.float32)
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0xFFF + ((u >> 13) & 1)) & 0xFFFFE000
        return u.astype(np.uint32).view(np.float32)
    if fmt == "bf16":
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32)
    raise ValueError(fmt)


# ------------------------------------------------------------ element lengths
def element_lengths(N, E, x, y, z, w, ogs_full):
    """(left, middle, right) lengths per direction, each [3][E] (x, y, z).

    ellipticMultiGridSchwarz.cpp:58-131 (harmonic mean of the distances
    between opposite faces over face-interior lines, weights w[j-1]*w[k-1] as
    written there) and :133-276 (neighbour lengths through an unmasked
    gather-scatter of face-interior values, fallback to the middle length)."""
    Nq = N + 1
    X = np.stack([np.asarray(c).reshape(E, Nq, Nq, Nq) for c in (x, y, z)])   # [c][e][k][j][i]
    if Nq == 2:
        rng = np.arange(0, 2)
        wt = np.ones((2, 2))
    else:
        rng = np.arange(1, Nq - 1)
        wt = np.outer(w[rng - 1], w[rng - 1])        # [a][b] = w[a-1] w[b-1]
    mid = np.zeros((3, E))
    # r: lines (k, j), weight w[j-1] w[k-1] -> wt[k][j]
    d = X[:, :, :, :, Nq - 1] - X[:, :, :, :, 0]        # [c][e][k][j]
    d2 = np.sum(d ** 2, axis=0)[:, rng][:, :, rng]      # [e][k][j]
    mid[0] = 1.0 / np.sqrt(np.sum(wt[None] / d2, axis=(1, 2)) / np.sum(wt))
    # s: lines (k, i), weight w[i-1] w[k-1] -> wt[k][i]
    d = X[:, :, :, Nq - 1, :] - X[:, :, :, 0, :]        # [c][e][k][i]
    d2 = np.sum(d ** 2, axis=0)[:, rng][:, :, rng]
    mid[1] = 1.0 / np.sqrt(np.sum(wt[None] / d2, axis=(1, 2)) / np.sum(wt))
    # t: lines (j, i), weight w[i-1] w[j-1] -> wt[j][i]
    d = X[:, :, Nq - 1, :, :] - X[:, :, 0, :, :]        # [c][e][j][i]
    d2 = np.sum(d ** 2, axis=0)[:, rng][:, :, rng]
    mid[2] = 1.0 / np.sqrt(np.sum(wt[None] / d2, axis=(1, 2)) / np.sum(wt))
    if np.any(np.abs(mid) < TOL_LEN) or np.any(mid < -TOL_LEN) or not np.all(np.isfinite(mid)):
        raise ValueError("element lengths: zero, negative or invalid length")
    if Nq == 2:
        return mid.copy(), mid, mid.copy()
    l = np.zeros((E, Nq, Nq, Nq))
    f = slice(1, Nq - 1)
    l[:, f, f, 0] = mid[0][:, None, None]
    l[:, f, f, Nq - 1] = mid[0][:, None, None]
    l[:, f, 0, f] = mid[1][:, None, None]
    l[:, f, Nq - 1, f] = mid[1][:, None, None]
    l[:, 0, f, f] = mid[2][:, None, None]
    l[:, Nq - 1, f, f] = mid[2][:, None, None]
    l = ogs_full.gs(l.ravel(), "add").reshape(E, Nq, Nq, Nq)
    left = np.stack([l[:, 1, 1, 0] - mid[0], l[:, 1, 0, 1] - mid[1], l[:, 0, 1, 1] - mid[2]])
    right = np.stack([l[:, 1, 1, Nq - 1] - mid[0], l[:, 1, Nq - 1, 1] - mid[1],
                      l[:, Nq - 1, 1, 1] - mid[2]])
    for a in (left, right):
        bad = (np.abs(a) < TOL_LEN) | (a < -TOL_LEN)
        a[bad] = mid[bad]
    return left, mid, right


def element_bcs(EToB, e):
    """(lbr, rbr, lbs, rbs, lbt, rbt): faces 4, 2, 1, 3, 0, 5 (:278-298)."""
    return tuple(int(EToB[6 * e + f]) for f in (4, 2, 1, 3, 0, 5))


# ------------------------------------------------------------ 1D operators
def stiffness_1d(D, w, lbc, rbc, ll, lm, lr):
    """compute_1d_stiffness_matrix (:308-378): nl = N + 3 points, A[row][col]."""
    n = len(w) - 1
    nl = n + 3
    ah = D.T @ (w[:, None] * D)          # ah[i][j] = sum_k D[k][i] w[k] D[k][j]
    a = np.zeros((nl, nl))
    i0 = 1 if lbc == 1 else 0
    i1 = n - 1 if rbc == 1 else n
    a[1, 1] = 1.0
    a[n + 1, n + 1] = 1.0
    fac = 2.0 / lm
    for j in range(i0, i1 + 1):
        for i in range(i0, i1 + 1):
            a[i + 1, j + 1] = fac * ah[i, j]
    if lbc == 0:
        fac = 2.0 / ll
        a[0, 0] = fac * ah[n - 1, n - 1]
        a[1, 0] = fac * ah[n, n - 1]
        a[0, 1] = fac * ah[n - 1, n]
        a[1, 1] += fac * ah[n, n]
    else:
        a[0, 0] = 1.0
    if rbc == 0:
        fac = 2.0 / lr
        a[n + 1, n + 1] += fac * ah[0, 0]
        a[n + 2, n + 1] = fac * ah[1, 0]
        a[n + 1, n + 2] = fac * ah[0, 1]
        a[n + 2, n + 2] = fac * ah[1, 1]
    else:
        a[n + 2, n + 2] = 1.0
    return a


def mass_1d(w, lbc, rbc, ll, lm, lr):
    """compute_1d_mass_matrix (:380-427): diagonal, nl = N + 3."""
    n = len(w) - 1
    nl = n + 3
    b = np.zeros((nl, nl))
    i0 = 1 if lbc == 1 else 0
    i1 = n - 1 if rbc == 1 else n
    b[1, 1] = 1.0
    b[n + 1, n + 1] = 1.0
    for i in range(i0, i1 + 1):
        b[i + 1, i + 1] = 0.5 * lm * w[i]
    if lbc == 0:
        b[0, 0] = 0.5 * ll * w[n - 1]
        b[1, 1] += 0.5 * ll * w[n]
    else:
        b[0, 0] = 1.0
    if rbc == 0:
        b[n + 1, n + 1] += 0.5 * lr * w[0]
        b[n + 2, n + 2] = 0.5 * lr * w[1]
    else:
        b[n + 2, n + 2] = 1.0
    return b


def matrices_1d(D, w, lbc, rbc, ll, lm, lr):
    """compute_1d_matrices (:505-556): (S[comp][eig], lam) with S^T B S = I,
    rows of the ghost / Dirichlet end points zeroed afterwards."""
    import scipy.linalg
    a = stiffness_1d(D, w, lbc, rbc, ll, lm, lr)
    b = mass_1d(w, lbc, rbc, ll, lm, lr)
    lam, S = scipy.linalg.eigh(a, b, type=1, lower=False, driver="gv")
    nl = a.shape[0]
    if lbc > 0:
        S[0, :] = 0.0
    if lbc == 1:
        S[1, :] = 0.0
    if rbc > 0:
        S[nl - 1, :] = 0.0
    if rbc == 1:
        S[nl - 2, :] = 0.0
    return S, lam


# ------------------------------------------------------------ extended mesh
def extended_mask(mesh_ext):
    """Masked ids of the order N+2 mesh (:672-743): edge nodes and nodes whose
    gs-min face flag is DIRICHLET."""
    m = mesh_ext
    Nq, E = m.Nq, m.E
    ogs = Ogs(m.gid)
    mapB = np.full((E, Nq, Nq, Nq), BIG, dtype=np.int32)
    faces = [(slice(None), 0), (1, 0), (2, Nq - 1), (1, Nq - 1), (2, 0), (0, Nq - 1)]
    for e in range(E):
        for f in range(6):
            bc = int(m.EToB[6 * e + f])
            if bc <= 0:
                continue
            v = mapB[e]
            if f == 0:
                v[0] = np.minimum(v[0], bc)
            elif f == 5:
                v[Nq - 1] = np.minimum(v[Nq - 1], bc)
            elif f == 1:
                v[:, 0] = np.minimum(v[:, 0], bc)
            elif f == 3:
                v[:, Nq - 1] = np.minimum(v[:, Nq - 1], bc)
            elif f == 4:
                v[:, :, 0] = np.minimum(v[:, :, 0], bc)
            elif f == 2:
                v[:, :, Nq - 1] = np.minimum(v[:, :, Nq - 1], bc)
    del faces
    mapB = ogs.gs(mapB.ravel(), "min").reshape(E, Nq, Nq, Nq)
    k, j, i = np.meshgrid(np.arange(Nq), np.arange(Nq), np.arange(Nq), indexing="ij")
    onb = lambda a: (a == 0) | (a == Nq - 1)                     # noqa: E731
    edge = (onb(i).astype(int) + onb(j).astype(int) + onb(k).astype(int)) >= 2
    masked = (mapB == DIRICHLET) | edge[None]
    return np.nonzero(masked.ravel())[0]


# ------------------------------------------------------------ kernels
def ghost_faces(nl):
    """(ghost layer, layer 2 from that face) slices of an [e][k][j][i] box."""
    f = slice(1, nl - 1)
    return [(np.s_[:, 0, f, f], np.s_[:, 2, f, f]), (np.s_[:, nl - 1, f, f], np.s_[:, nl - 3, f, f]),
            (np.s_[:, f, 0, f], np.s_[:, f, 2, f]), (np.s_[:, f, nl - 1, f], np.s_[:, f, nl - 3, f]),
            (np.s_[:, f, f, 0], np.s_[:, f, f, 2]), (np.s_[:, f, f, nl - 1], np.s_[:, f, f, nl - 3])]


def pre_fdm(u, N):
    """preFDM.c:3-98: interior = u, face ghost layers = the element's own
    second layer, edges and corners 0."""
    Nq, nl = N + 1, N + 3
    u = np.asarray(u).reshape(-1, Nq, Nq, Nq)
    w = np.zeros((u.shape[0], nl, nl, nl), dtype=u.dtype)
    f = slice(1, nl - 1)
    w[:, f, f, f] = u
    w[:, 0, f, f] = u[:, 1]
    w[:, nl - 1, f, f] = u[:, Nq - 2]
    w[:, f, 0, f] = u[:, :, 1]
    w[:, f, nl - 1, f] = u[:, :, Nq - 2]
    w[:, f, f, 0] = u[:, :, :, 1]
    w[:, f, f, nl - 1] = u[:, :, :, Nq - 2]
    return w


def fdm_box(w, S, invL, T=np.float32, fmt=None):
    """(Sz x Sy x Sx) invL (Sz x Sy x Sx)^T on each box (fusedFDM.c:121-175).
    S: [3][E][nl][nl] as S[d][e][component][eigenvector]; invL: [E][k][j][i].
    Large inputs are split into element chunks handled by a thread pool (numpy
    releases the GIL in these loops); every element is computed alone, so the
    result does not depend on the split."""
    E = w.shape[0]
    if E >= 1024:
        import concurrent.futures as cf
        nt = max(1, min(16, os.cpu_count() or 1))
        cuts = np.linspace(0, E, nt + 1).astype(int)
        out = np.empty(w.shape, dtype=T)

        def part(a, b):
            out[a:b] = _fdm_box(w[a:b], S[:, a:b], invL[a:b], T, fmt)
        with cf.ThreadPoolExecutor(nt) as ex:
            list(ex.map(lambda ab: part(*ab), zip(cuts[:-1], cuts[1:])))
        return out
    return _fdm_box(w, S, invL, T, fmt)


def _fdm_box(w, S, invL, T, fmt):
    if fmt is None:
        Sx, Sy, Sz = S[0], S[1], S[2]
        t = np.einsum("elm,ekjl->ekjm", Sx, w)
        t = np.einsum("eln,eklm->eknm", Sy, t)
        t = np.einsum("elp,elnm->epnm", Sz, t)
        t = (t * invL).astype(T)
        t = np.einsum("eim,epnm->epni", Sx, t)
        t = np.einsum("ejn,epni->epji", Sy, t)
        return np.einsum("ekp,epji->ekji", Sz, t).astype(T)
    Sr = [round_fmt(np.asarray(S[d], dtype=np.float32), fmt, operator=True).astype(np.float64)
          for d in range(3)]
    r = lambda a: round_fmt(a.astype(np.float32), fmt).astype(np.float64)   # noqa: E731
    f32 = lambda a: a.astype(np.float32)                                      # noqa: E731
    t = f32(np.einsum("elm,ekjl->ekjm", Sr[0], r(w)))
    t = f32(np.einsum("eln,eklm->eknm", Sr[1], r(t)))
    t = f32(np.einsum("elp,elnm->epnm", Sr[2], r(t)))
    t = f32(t * np.asarray(invL, dtype=np.float32))
    t = f32(np.einsum("eim,epnm->epni", Sr[0], r(t)))
    t = f32(np.einsum("ejn,epni->epji", Sr[1], r(t)))
    return f32(np.einsum("ekp,epji->ekji", Sr[2], r(t)))


def fused_fdm(w, S, invL, T=np.float32, fmt=None):
    """fusedFDM.c:47-175 on gathered boxes: ghost -= own layer 2 (leaving the
    neighbour's overlap value), then the FDM solve; returns the full box."""
    w = np.array(w, dtype=T, copy=True)
    for g, inner in ghost_faces(w.shape[1]):
        w[g] = w[g] - w[inner]
    return fdm_box(w, S, invL, T, fmt)


def post_fdm(x, own, wts, N, T=np.float32):
    """postFDM.c:3-155: x = gathered FDM boxes, own = this element's boxes;
    ghost -= own ghost, layer 2 += ghost (z faces, then y, then x), restrict
    to the Nq^3 interior and scale by wts."""
    nl = N + 3
    x = np.array(x, dtype=T, copy=True).reshape(-1, nl, nl, nl)
    own = np.asarray(own).reshape(x.shape)
    for g, _ in ghost_faces(nl):
        x[g] = x[g] - own[g]
    f = slice(1, nl - 1)
    x[:, 2, f, f] += x[:, 0, f, f]
    x[:, nl - 3, f, f] += x[:, nl - 1, f, f]
    x[:, f, 2, f] += x[:, f, 0, f]
    x[:, f, nl - 3, f] += x[:, f, nl - 1, f]
    x[:, f, f, 2] += x[:, f, f, 0]
    x[:, f, f, nl - 3] += x[:, f, f, nl - 1]
    return (x[:, f, f, f].ravel() * np.asarray(wts, dtype=T)).astype(T)


# ------------------------------------------------------------ the smoother
class Schwarz:
    """FDM operators and the ASM / RAS application on one multigrid level.

    level_mesh: oracle BoxMesh of the level (order N, ids, EToB, D);
    fine: (N, x, y, z, w, ogs_full) of the finest level, whose coordinates
    give the element lengths (pMGLevel::build passes the finest elliptic);
    maskIds, ogs (masked), invDegree of the level."""

    def __init__(self, level_mesh, fine, maskIds, ogs, invDegree, mode="asm", T=np.float32,
                 fmt=None, mesh_ext=None, poisson=True, lambda0=1.0, lambda1=0.0):
        m = level_mesh
        self.mode, self.T, self.fmt = mode, T, fmt
        self.N, self.E = m.N, m.E
        self.Nq, self.nl = m.N + 1, m.N + 3
        self.maskIds, self.ogs = maskIds, ogs
        self.invDegreeT = invDegree.astype(T)
        Nf, xf, yf, zf, wf, ogs_full_f = fine
        self.left, self.mid, self.right = element_lengths(Nf, m.E, xf, yf, zf, wf, ogs_full_f)
        r, w = gll(m.N)
        D = m.D
        nl = self.nl
        S = np.zeros((3, m.E, nl, nl))
        lam = np.zeros((3, m.E, nl))
        for e in range(m.E):
            bc = element_bcs(m.EToB, e)
            for d in range(3):
                S[d, e], lam[d, e] = matrices_1d(D, w, bc[2 * d], bc[2 * d + 1],
                                                 self.left[d, e], self.mid[d, e], self.right[d, e])
        # Helmholtz (an extension: the reference's Schwarz smoother is
        # Poisson-only, ellipticSetup.cpp:59-64): invert lambda0 (Lx + Ly + Lz)
        # + lambda1, the mass matrix being the identity in the eigenbasis; the
        # modes with Lx + Ly + Lz <= 1e-5 (constant on a box with free ends)
        # stay dropped as in the Poisson case
        l0, l1 = (1.0, 0.0) if poisson else (lambda0, lambda1)
        self.S64, self.lam = S, lam
        diag = lam[2][:, :, None, None] + lam[1][:, None, :, None] + lam[0][:, None, None, :]
        with np.errstate(divide="ignore"):
            invL = np.where(diag > EPS_DIAG, 1.0 / (l0 * diag + l1), 0.0)
        self.invL = invL.astype(T)                          # [e][k][j][i]
        self.S = S.astype(T)
        # extended mesh: ids from an order N+2 box of the same elements (or the
        # caller's order N+2 mesh with gid and EToB, for reoriented elements)
        me = (BoxMesh(m.nx, m.ny, m.nz, m.N + 2, bc=getattr(m, "bc", "dirichlet"),
                      periodic=getattr(m, "periodic", "")) if mesh_ext is None else mesh_ext)
        mids = extended_mask(me)
        g = me.gid.copy()
        g[mids] = 0
        self.ogs_ext = Ogs(g)
        if mode == "asm":
            self.wts = self._weights()

    def apply(self, u):
        """smoothSchwarz(u, Su) (:1056-1156), masked output."""
        T = self.T
        E, nl = self.E, self.nl
        w = pre_fdm(np.asarray(u, dtype=T), self.N)
        w = self.ogs_ext.gs(w.ravel(), "add").reshape(E, nl, nl, nl)
        out = fused_fdm(w, self.S, self.invL, T, self.fmt)
        f = slice(1, nl - 1)
        if self.mode == "ras":
            su = (out[:, f, f, f].ravel() * self.invDegreeT).astype(T)
        else:
            x = self.ogs_ext.gs(out.ravel(), "add")
            su = post_fdm(x, out, self.wts, self.N, T)
        su = self.ogs.gs(su, "add")
        su[self.maskIds] = 0
        return su

    def _weights(self):
        """generate_weights (:798-835): 1 / (gs of 1 + overlap counts)."""
        T = self.T
        E, nl = self.E, self.nl
        f = slice(1, nl - 1)
        w1 = np.ones((E, nl, nl, nl), dtype=T)
        w2 = np.ones((E, nl, nl, nl), dtype=T)
        w1 = self.ogs_ext.gs(w1.ravel(), "add").reshape(E, nl, nl, nl)
        # extrude(work1, 0, 1, work2, 0, -1), then extrude(work1, 2, 1, work1, 0, 1): x, y, z
        for g in (np.s_[:, f, f, 0], np.s_[:, f, f, nl - 1], np.s_[:, f, 0, f], np.s_[:, f, nl - 1, f],
                  np.s_[:, 0, f, f], np.s_[:, nl - 1, f, f]):
            w1[g] = w1[g] - w2[g]
        w1[:, f, f, 2] += w1[:, f, f, 0]
        w1[:, f, f, nl - 3] += w1[:, f, f, nl - 1]
        w1[:, f, 2, f] += w1[:, f, 0, f]
        w1[:, f, nl - 3, f] += w1[:, f, nl - 1, f]
        w1[:, 2, f, f] += w1[:, 0, f, f]
        w1[:, nl - 3, f, f] += w1[:, nl - 1, f, f]
        wts = self.ogs.gs(w1[:, f, f, f].ravel().astype(T), "add")
        return (T(1) / wts).astype(T)
