"""CPU oracle for the nekRS pressure-Poisson hot path.

TEST INFRASTRUCTURE ONLY.  This package is the parity checker: only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may import it.
The product (``nekrs_amd``) never imports it and has no CPU fallback.

``liboracle.so`` is built from ``nrs_oracle.c`` (a C restatement of the
reference's serial kernels and solver driver; each function cites the
reference file:line it follows).  ``_ref/libnrsref.so`` is built by
``build_ref.sh`` from the reference's own serial kernels and gslib, only where
the reference tree exists (the build container); it never travels.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnrsref.so")

OP = {"add": 0, "mul": 1, "min": 2, "max": 3}
MAP = {"none": 0, "sine": 1, "kershaw": 2}
G00, G01, G11, G12, G02, G22, GWJ = range(7)
DIRICHLET = 1

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force=False):
    """Compile liboracle.so with gcc (OpenMP)."""
    src = os.path.join(HERE, "nrs_oracle.c")
    if not force and os.path.exists(LIB_PATH) and os.path.getmtime(LIB_PATH) >= os.path.getmtime(src):
        return LIB_PATH
    subprocess.check_call(["gcc", "-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared",
                           "-o", LIB_PATH + ".tmp", src, "-lm"])
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def build_ref(reference_root="/root/reference"):
    """Build oracle/_ref/libnrsref.so from the reference sources (build container only)."""
    if not os.path.isdir(os.path.join(reference_root, "kernels", "elliptic")):
        return None
    subprocess.check_call(["bash", os.path.join(HERE, "build_ref.sh"), reference_root],
                          stdout=subprocess.DEVNULL)
    return REF_PATH


class _OrcOgs(C.Structure):
    _fields_ = [("N", C.c_int64), ("Nrows", C.c_int64), ("Nids", C.c_int64),
                ("offsets", C.POINTER(C.c_int64)), ("ids", C.POINTER(C.c_int64)),
                ("invDegree", C.POINTER(C.c_double))]


class _OrcElliptic(C.Structure):
    _fields_ = [("N", C.c_int), ("E", C.c_int),
                ("ggeo", C.POINTER(C.c_double)), ("D", C.POINTER(C.c_double)),
                ("ogs", C.POINTER(_OrcOgs)), ("Nmasked", C.c_int64),
                ("maskIds", C.POINTER(C.c_int64)),
                ("lambda0", C.c_double), ("lambda1", C.c_double),
                ("poisson", C.c_int), ("volume", C.c_double), ("allNeumann", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_gll.argtypes = [C.c_int, _dp, _dp]
        L.orc_dmatrix.argtypes = [C.c_int, _dp, _dp]
        L.orc_interp.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_kershaw_map.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_double)] * 3
        L.orc_box_mesh.argtypes = [C.c_int] * 5 + [C.c_double, _dp, _dp, _dp, _i64p, C.c_void_p]
        L.orc_geofacs.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_void_p]
        L.orc_ogs_setup.argtypes = [C.c_int64, _i64p]
        L.orc_ogs_setup.restype = C.POINTER(_OrcOgs)
        L.orc_ogs_free.argtypes = [C.POINTER(_OrcOgs)]
        L.orc_ogs_get.argtypes = [C.POINTER(_OrcOgs), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_gs_f64.argtypes = [C.POINTER(_OrcOgs), _dp, C.c_int, C.c_int64, C.c_int]
        L.orc_gs_f32.argtypes = [C.POINTER(_OrcOgs), _fp, C.c_int, C.c_int64, C.c_int]
        L.orc_gs_i32.argtypes = [C.POINTER(_OrcOgs), _i32p, C.c_int, C.c_int64, C.c_int]
        L.orc_mask_ids.argtypes = [C.c_int, C.c_int, C.POINTER(_OrcOgs), _i32p, C.c_void_p]
        L.orc_mask_ids.restype = C.c_int64
        L.orc_mask_f64.argtypes = [C.c_int64, _i64p, _dp]
        L.orc_ax_f64.argtypes = [C.c_int, C.c_int, C.c_void_p, _dp, _dp, _dp, _dp, C.c_int,
                                 C.c_int, _dp, _dp]
        L.orc_ax_f32.argtypes = [C.c_int, C.c_int, C.c_void_p, _fp, _fp, _fp, _fp, C.c_int,
                                 C.c_int, _fp, _fp]
        L.orc_block_ax_f64.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64, _dp,
                                       _dp, _dp, _dp, _dp, _dp]
        L.orc_build_diag_f64.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_int,
                                         C.c_int, _dp]
        L.orc_coarsen.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_prolongate.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_wdot.argtypes = [C.c_int64, _dp, _dp, _dp]
        L.orc_wdot.restype = C.c_double
        L.orc_update_pcg.argtypes = [C.c_int64, _dp, _dp, C.c_double, _dp]
        L.orc_update_pcg.restype = C.c_double
        L.orc_operator.argtypes = [C.POINTER(_OrcElliptic), _dp, _dp]
        L.orc_pcg.argtypes = [C.POINTER(_OrcElliptic), C.c_int, C.c_void_p, C.c_int, _dp, _dp,
                              C.c_double, C.c_int, C.POINTER(C.c_double), C.c_void_p]
        L.orc_solve.argtypes = [C.POINTER(_OrcElliptic), C.c_int, C.c_void_p, C.c_int, C.c_double,
                                C.c_int, C.c_int, _dp, _dp, C.POINTER(C.c_double),
                                C.POINTER(C.c_double), C.c_void_p]
        _lib = L
    return _lib


# ---------------------------------------------------------------- 1D basis
def gll(N):
    z = np.zeros(N + 1); w = np.zeros(N + 1)
    lib().orc_gll(N, z, w)
    return z, w


def dmatrix(N, z=None):
    if z is None:
        z = gll(N)[0]
    D = np.zeros((N + 1) * (N + 1))
    lib().orc_dmatrix(N, np.ascontiguousarray(z, dtype=np.float64), D)
    return D.reshape(N + 1, N + 1)


def interp(zin, zout):
    zin = np.ascontiguousarray(zin, dtype=np.float64)
    zout = np.ascontiguousarray(zout, dtype=np.float64)
    I = np.zeros(len(zout) * len(zin))
    lib().orc_interp(len(zin), zin, len(zout), zout, I)
    return I.reshape(len(zout), len(zin))


def kershaw_map(eps, x, y, z):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_kershaw_map(eps, eps, x, y, z, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


# ---------------------------------------------------------------- mesh
NEUMANN = 2


class BoxMesh:
    """Structured box [0,1]^3 (lexicographic elements, GLL nodes).

    bc: "dirichlet" (DIRICHLET = 1 on every boundary face) or "neumann"
    (NEUMANN = 2, a closed box); periodic: the axes ("x", "y", "z") whose
    opposite faces are connected (the last node layer takes the ids of the
    first, those faces carry no boundary type)."""

    def __init__(self, nx, ny, nz, N, map="none", eps=0.0, bc="dirichlet", periodic=""):
        assert bc in ("dirichlet", "neumann")
        self.nx, self.ny, self.nz, self.N = nx, ny, nz, N
        self.bc, self.periodic = bc, "".join(sorted(periodic))
        self.Nq = N + 1
        self.Np = self.Nq ** 3
        self.E = nx * ny * nz
        L = self.E * self.Np
        self.x = np.zeros(L); self.y = np.zeros(L); self.z = np.zeros(L)
        self.gid = np.zeros(L, dtype=np.int64)
        self.EToB = np.zeros(self.E * 6, dtype=np.int32)
        lib().orc_box_mesh(nx, ny, nz, N, MAP[map], eps, self.x, self.y, self.z, self.gid,
                           self.EToB.ctypes.data)
        if bc == "neumann":
            self.EToB[self.EToB == DIRICHLET] = NEUMANN
        if self.periodic:
            gx, gy = nx * N + 1, ny * N + 1
            g = self.gid - 1
            G = [g % gx, (g // gx) % gy, g // (gx * gy)]
            for a, (ax, n) in enumerate(zip("xyz", (nx, ny, nz))):
                if ax in self.periodic:
                    G[a] = G[a] % (n * N)
            self.gid = 1 + G[0] + gx * (G[1] + gy * G[2])
            eb = self.EToB.reshape(-1, 6)
            for ax, faces in (("x", (2, 4)), ("y", (1, 3)), ("z", (0, 5))):
                if ax in self.periodic:
                    eb[:, list(faces)] = 0
        self.r, self.w = gll(N)
        self.D = dmatrix(N, self.r)
        self.ggeo = np.zeros(self.E * 7 * self.Np)
        self.J = np.zeros(L)
        lib().orc_geofacs(N, self.E, self.D.ravel().copy(), self.w, self.x, self.y, self.z,
                          self.ggeo, self.J.ctypes.data)

    @property
    def Nlocal(self):
        return self.E * self.Np


class ArrayMesh:
    """A mesh given by its GLL coordinates, global ids and boundary flags (for
    meshes read from files, e.g. with curved elements): the same fields and
    geometric factors as BoxMesh."""

    def __init__(self, N, x, y, z, gid, EToB):
        self.N, self.Nq = N, N + 1
        self.Np = self.Nq ** 3
        self.x, self.y, self.z = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (x, y, z))
        self.E = self.x.size // self.Np
        self.gid = np.ascontiguousarray(gid, dtype=np.int64).copy()
        self.EToB = np.ascontiguousarray(EToB, dtype=np.int32).copy()
        self.r, self.w = gll(N)
        self.D = dmatrix(N, self.r)
        self.ggeo = np.zeros(self.E * 7 * self.Np)
        self.J = np.zeros(self.E * self.Np)
        lib().orc_geofacs(N, self.E, self.D.ravel().copy(), self.w, self.x, self.y, self.z,
                          self.ggeo, self.J.ctypes.data)

    @property
    def Nlocal(self):
        return self.E * self.Np


# ---------------------------------------------------------------- ogs
class Ogs:
    """Local gather-scatter handle (restates ogsSetup's local rows)."""

    def __init__(self, gids):
        gids = np.ascontiguousarray(gids, dtype=np.int64)
        self._p = lib().orc_ogs_setup(len(gids), gids)
        o = self._p.contents
        self.N, self.Nrows, self.Nids = o.N, o.Nrows, o.Nids
        self.offsets = np.zeros(self.Nrows + 1, dtype=np.int64)
        self.ids = np.zeros(max(self.Nids, 1), dtype=np.int64)
        self.invDegree = np.zeros(max(self.N, 1))
        lib().orc_ogs_get(self._p, self.offsets.ctypes.data, self.ids.ctypes.data,
                          self.invDegree.ctypes.data)
        self.ids = self.ids[: self.Nids]
        self.invDegree = self.invDegree[: self.N]

    def __del__(self):
        try:
            lib().orc_ogs_free(self._p)
        except Exception:
            pass

    def gs(self, v, op="add", nvec=1, stride=None):
        v = np.ascontiguousarray(v).copy()
        stride = self.N if stride is None else stride
        fn = {np.float64: lib().orc_gs_f64, np.float32: lib().orc_gs_f32,
              np.int32: lib().orc_gs_i32}[v.dtype.type]
        fn(self._p, v, nvec, stride, OP[op])
        return v


def mask_ids(mesh, ogs_unmasked):
    n = lib().orc_mask_ids(mesh.N, mesh.E, ogs_unmasked._p, mesh.EToB, None)
    ids = np.zeros(max(n, 1), dtype=np.int64)
    lib().orc_mask_ids(mesh.N, mesh.E, ogs_unmasked._p, mesh.EToB, ids.ctypes.data)
    return ids[:n]


def masked_gids(gid, maskIds):
    g = gid.copy()
    g[maskIds] = 0
    return g


# ---------------------------------------------------------------- kernels
def ax(N, ggeo, D, q, lambda0=1.0, lambda1=0.0, poisson=True, elist=None, dtype=np.float64):
    """Element-local Ax (no mask, no gs).  lambda0/lambda1: scalars or per-node fields."""
    q = np.ascontiguousarray(q, dtype=dtype)
    E = q.size // (N + 1) ** 3
    field = not np.isscalar(lambda0)
    l0 = np.ascontiguousarray(np.atleast_1d(lambda0), dtype=dtype)
    l1 = np.ascontiguousarray(np.atleast_1d(lambda1), dtype=dtype)
    Aq = np.zeros_like(q)
    el = None
    if elist is not None:
        el = np.ascontiguousarray(elist, dtype=np.int32)
        E = len(el)
    fn = lib().orc_ax_f64 if dtype == np.float64 else lib().orc_ax_f32
    fn(N, E, el.ctypes.data if el is not None else None,
       np.ascontiguousarray(ggeo, dtype=dtype), np.ascontiguousarray(D, dtype=dtype).ravel(),
       l0, l1, int(poisson), int(field), q, Aq)
    return Aq


def build_diag(N, ggeo, D, lambda0=1.0, lambda1=0.0, poisson=True, E=None):
    Np = (N + 1) ** 3
    E = len(ggeo) // (7 * Np) if E is None else E
    field = not np.isscalar(lambda0)
    diag = np.zeros(E * Np)
    lib().orc_build_diag_f64(N, E, np.ascontiguousarray(ggeo), np.ascontiguousarray(D).ravel(),
                             np.ascontiguousarray(np.atleast_1d(lambda0), dtype=np.float64),
                             np.ascontiguousarray(np.atleast_1d(lambda1), dtype=np.float64),
                             int(poisson), int(field), diag)
    return diag


def transfer_R(Nf, Nc):
    """R = transpose(InterpolationMatrix1D(Nc -> Nf)), shape [NqC][NqF]."""
    zf, _ = gll(Nf)
    zc, _ = gll(Nc)
    return interp(zc, zf).T.copy()


def coarsen(Nf, Nc, R, qf):
    E = qf.size // (Nf + 1) ** 3
    qc = np.zeros(E * (Nc + 1) ** 3)
    lib().orc_coarsen(Nf, Nc, E, np.ascontiguousarray(R).ravel(), np.ascontiguousarray(qf), qc)
    return qc


def prolongate(Nf, Nc, R, qc, qf):
    E = qc.size // (Nc + 1) ** 3
    out = np.ascontiguousarray(qf, dtype=np.float64).copy()
    lib().orc_prolongate(Nf, Nc, E, np.ascontiguousarray(R).ravel(), np.ascontiguousarray(qc), out)
    return out


def wdot(w, a, b):
    return lib().orc_wdot(len(a), np.ascontiguousarray(w), np.ascontiguousarray(a),
                          np.ascontiguousarray(b))


def update_pcg(w, Ap, alpha, r):
    """r - alpha Ap and sum w r^2 (ellipticBlockUpdatePCG)."""
    r = np.ascontiguousarray(r, dtype=np.float64).copy()
    rr = lib().orc_update_pcg(len(r), np.ascontiguousarray(w), np.ascontiguousarray(Ap), alpha, r)
    return r, rr


# ---------------------------------------------------------------- solver
class Elliptic:
    """Poisson/Helmholtz operator on a box mesh with Dirichlet faces.

    Mirrors the state ellipticSolveSetup builds: masked ogs, maskIds,
    invDegree (from the masked ogs), constant lambda0/lambda1."""

    def __init__(self, mesh, lambda0=1.0, lambda1=0.0, poisson=True, zero_mean=None):
        """all_neumann: Poisson with no face of a boundary type other than
        NEUMANN (ellipticSetup.cpp:187-206); zero_mean (default: all_neumann)
        makes solve() remove the mean of r, z and x (ellipticZeroMean)."""
        self.mesh = mesh
        eb = np.asarray(mesh.EToB)
        self.all_neumann = bool(poisson) and not np.any((eb > 0) & (eb != NEUMANN))
        self.zero_mean = self.all_neumann if zero_mean is None else bool(zero_mean)
        self.ogs_full = Ogs(mesh.gid)
        self.maskIds = mask_ids(mesh, self.ogs_full)
        self.ogs = Ogs(masked_gids(mesh.gid, self.maskIds))
        self.invDegree = self.ogs.invDegree
        self.lambda0, self.lambda1, self.poisson = lambda0, lambda1, poisson
        self.volume = float(np.sum(mesh.ggeo.reshape(mesh.E, 7, mesh.Np)[:, GWJ, :]))
        self._D = np.ascontiguousarray(mesh.D).ravel().copy()
        self._c = _OrcElliptic(mesh.N, mesh.E, mesh.ggeo.ctypes.data_as(C.POINTER(C.c_double)),
                               self._D.ctypes.data_as(C.POINTER(C.c_double)), self.ogs._p,
                               len(self.maskIds),
                               self.maskIds.ctypes.data_as(C.POINTER(C.c_int64)),
                               lambda0, lambda1, int(poisson), self.volume,
                               int(self.zero_mean))

    def operator(self, q):
        Aq = np.zeros(self.mesh.Nlocal)
        lib().orc_operator(C.byref(self._c), np.ascontiguousarray(q, dtype=np.float64), Aq)
        return Aq

    def inv_diag(self):
        m = self.mesh
        d = build_diag(m.N, m.ggeo, m.D, self.lambda0, self.lambda1, self.poisson)
        d = self.ogs.gs(d, "add")
        return 1.0 / d

    def solve(self, rhs, x0=None, precon="jacobi", flexible=False, tol=1e-8, relative=True,
              maxit=500):
        m = self.mesh
        x = np.zeros(m.Nlocal) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        inv = self.inv_diag() if precon == "jacobi" else None
        res0, res = C.c_double(), C.c_double()
        hist = np.zeros(maxit + 1)
        it = lib().orc_solve(C.byref(self._c), 1 if precon == "jacobi" else 0,
                             inv.ctypes.data if inv is not None else None, int(flexible), tol,
                             int(relative), maxit, np.ascontiguousarray(rhs, dtype=np.float64), x,
                             C.byref(res0), C.byref(res), hist.ctypes.data)
        return x, {"iterations": it, "res0": res0.value, "res": res.value,
                   "history": hist[: abs(it)].copy()}


class FieldElliptic(Elliptic):
    """The same state with per-node coefficients (ELLIPTIC COEFF FIELD =
    TRUE): ellipticOperator with o_lambda0 / o_lambda1 fields
    (ellipticOperator.cpp:77-99: Ax with the field kernel, mask, gather-scatter)
    and the Jacobi diagonal of ellipticUpdateJacobi (ellipticUpdateJacobi.cpp:
    62-85: element diagonal with the fields, gather-scatter, reciprocal).
    Built from ax() and build_diag(), which take fields; the solve is
    oracle.mg.fcg_solve (pcg) with this operator."""

    def __init__(self, mesh, lambda0_field, lambda1_field=None, poisson=True):
        super().__init__(mesh, 1.0, 0.0, poisson)
        self.lambda0 = np.ascontiguousarray(lambda0_field, dtype=np.float64)
        self.lambda1 = (np.ascontiguousarray(lambda1_field, dtype=np.float64) if lambda1_field is not None
                        else np.zeros(mesh.Nlocal))

    def operator(self, q):
        m = self.mesh
        Aq = ax(m.N, m.ggeo, m.D, q, self.lambda0, self.lambda1, self.poisson)
        Aq[self.maskIds] = 0.0
        return self.ogs.gs(Aq, "add")

    def inv_diag(self):
        m = self.mesh
        d = build_diag(m.N, m.ggeo, m.D, self.lambda0, self.lambda1, self.poisson)
        return 1.0 / self.ogs.gs(d, "add")


# ---------------------------------------------------------------- reference build
_ref = None


def ref_lib():
    """The reference's own serial kernels + gslib (None when not built)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            return None
        _ref = C.CDLL(REF_PATH)
    return _ref
