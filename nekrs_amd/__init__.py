"""nekrs_amd -- MI355X (gfx950) pressure-Poisson hot path of nekRS.

Python mirror of the reference's solver interface over the C ABI of
libnekrs_hip.so (include/nrs_hip.h).  Device memory is torch tensors on
``cuda`` (ROCm); their data pointers and the current HIP stream are handed to
the library.  All arithmetic runs in the library's HIP kernels.

  ogs = Ogs(global_ids)                  # ogsSetup
  ogs.gather_scatter(v, "add")           # oogs::startFinish
  ax(N, ggeo, D, q, Aq)                  # ellipticPartialAxCoeffHex3D
  el = Elliptic(mesh, precon="jacobi")   # ellipticSolveSetup
  el.operator(q, Aq)                     # ellipticOperator
  stats = el.solve(rhs, x)               # ellipticSolve
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import NrsError

__all__ = ["NrsError", "init", "parse_mg_schedule", "box_mesh", "gll", "ax", "block_ax", "build_diagonal", "mask",
           "re2_curves", "mesh_midside",
           "coarsen", "prolongate", "geometric_factors", "Ogs", "Elliptic", "BoxMesh", "fdm_1d",
           "fdm_element_lengths", "Batch", "make_desc", "make_opts", "tuned_variant", "set_variant",
           "PressureRhs", "BC", "read_re2", "mesh_from_vertices", "Re2Mesh", "linalg_reduce",
           "block_update_pcg"]

OPS = {"add": _lib.ADD, "mul": _lib.MUL, "min": _lib.MIN, "max": _lib.MAX}
MAPS = {"none": 0, "sine": 1, "kershaw": 2}
SOLVERS = {"pcg": 0, "fcg": 1, "pcg+flexible": 1, "gmres": 2, "pgmres": 2, "fgmres": 3,
           "pgmres+flexible": 3}
PRECONS = {"none": 0, "jacobi": 1, "multigrid": 2, "user": 3}
SMOOTHERS = {"chebyshev+jacobi": 0, "opt4chebyshev+jacobi": 1, "chebyshev+asm": 2,
             "opt4chebyshev+asm": 3, "asm": 4, "ras": 5, "chebyshev+ras": 6,
             "opt4chebyshev+ras": 7}
FDM_PRECISIONS = {"f32": 0, "bf16": 1, "fp16": 2}
DTYPES = {"f64": _lib.F64, "f32": _lib.F32}


def _torch():
    import torch
    return torch


def _dtype_code(t):
    torch = _torch()
    if t.dtype == torch.float64:
        return _lib.F64
    if t.dtype == torch.float32:
        return _lib.F32
    if t.dtype == torch.int32:
        return _lib.I32
    raise TypeError("unsupported dtype %s" % t.dtype)


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _stream(stream=None):
    torch = _torch()
    s = torch.cuda.current_stream() if stream is None else stream
    return C.c_void_p(s.cuda_stream)


class _DeviceArray:
    """A device buffer owned by the library, for torch.as_tensor (CUDA array
    interface; torch takes no read-only views, so r is writable but must not be
    written)."""

    def __init__(self, addr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "version": 2,
                                         "data": (int(addr), False), "strides": None}


def _device_view(addr, n, typestr="<f8"):
    return _torch().as_tensor(_DeviceArray(addr, n, typestr), device="cuda")


def init(device=0):
    _lib.call("nrs_init", device)


# ------------------------------------------------------------------ setup helpers
def gll(N):
    z = np.zeros(N + 1); w = np.zeros(N + 1); D = np.zeros((N + 1) * (N + 1))
    _lib.call("nrs_gll", N, z.ctypes.data, w.ctypes.data, D.ctypes.data)
    return z, w, D.reshape(N + 1, N + 1)


def box_mesh(nx, ny, nz, N, map="none", eps=0.0):
    """Host arrays (x, y, z, globalIds, EToB) of a structured box [0,1]^3."""
    E = nx * ny * nz
    L = E * (N + 1) ** 3
    x = np.zeros(L); y = np.zeros(L); z = np.zeros(L)
    gid = np.zeros(L, dtype=np.int64)
    EToB = np.zeros(E * 6, dtype=np.int32)
    _lib.call("nrs_box_mesh", nx, ny, nz, N, MAPS[map], float(eps), x.ctypes.data, y.ctypes.data,
              z.ctypes.data, gid.ctypes.data, EToB.ctypes.data)
    return x, y, z, gid, EToB


class BoxMesh:
    """Host-side description of a box mesh (coordinates, global ids, BCs)."""

    def __init__(self, nx, ny, nz, N, map="none", eps=0.0):
        self.nx, self.ny, self.nz, self.N = nx, ny, nz, N
        self.Nq = N + 1
        self.Np = self.Nq ** 3
        self.E = nx * ny * nz
        self.Nlocal = self.E * self.Np
        self.map, self.eps = map, eps
        self.x, self.y, self.z, self.gid, self.EToB = box_mesh(nx, ny, nz, N, map, eps)


def read_re2(path):
    """Corners [E][8] (lexicographic), groups [E] and boundary codes [E][6] (the
    solver's face order) of a .re2 file, plus its header counts."""
    path_b = str(path).encode()
    nelgt, nelgv, ncurve, nbc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    _lib.call("nrs_re2_info", path_b, C.byref(nelgt), C.byref(nelgv), C.byref(ncurve), C.byref(nbc))
    E = nelgt.value
    vx, vy, vz = (np.zeros(E * 8) for _ in range(3))
    group = np.zeros(E, dtype=np.int32)
    cbc = np.zeros(E * 6 * 3, dtype=np.uint8)
    _lib.call("nrs_re2_read", path_b, vx.ctypes.data, vy.ctypes.data, vz.ctypes.data,
              group.ctypes.data, cbc.ctypes.data)
    codes = np.array([bytes(cbc[3 * n:3 * n + 3]).decode("ascii", "replace")
                      for n in range(E * 6)]).reshape(E, 6)
    peer = np.zeros(E * 12, dtype=np.int32)
    _lib.call("nrs_re2_periodic", path_b, peer.ctypes.data)
    return {"nelgt": E, "nelgv": nelgv.value, "ncurve": ncurve.value, "nbc": nbc.value,
            "vx": vx, "vy": vy, "vz": vz, "group": group, "cbc": codes,
            "peer": peer.reshape(E, 6, 2)}


def mesh_from_vertices(N, vx, vy, vz, tol=0.0, peer=None):
    """(x, y, z, globalIds) of the order-N GLL nodes of a conforming hex mesh
    from its corners [E*8] (lexicographic), nrs_mesh_from_vertices; with peer
    ([E][6][2] periodic face pairs, -1 elsewhere) the paired faces are connected
    (nrs_mesh_from_vertices_periodic)."""
    vx, vy, vz = (np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy, vz))
    E = vx.size // 8
    L = E * (N + 1) ** 3
    x, y, z = np.zeros(L), np.zeros(L), np.zeros(L)
    gid = np.zeros(L, dtype=np.int64)
    if peer is None:
        _lib.call("nrs_mesh_from_vertices", N, E, vx.ctypes.data, vy.ctypes.data, vz.ctypes.data,
                  float(tol), x.ctypes.data, y.ctypes.data, z.ctypes.data, gid.ctypes.data)
    else:
        pr = np.ascontiguousarray(peer, dtype=np.int32).ravel()
        assert pr.size == E * 12
        _lib.call("nrs_mesh_from_vertices_periodic", N, E, vx.ctypes.data, vy.ctypes.data,
                  vz.ctypes.data, float(tol), pr.ctypes.data, x.ctypes.data, y.ctypes.data,
                  z.ctypes.data, gid.ctypes.data)
    return x, y, z, gid


def re2_curves(path):
    """Curved-side records of a .re2 file (nrs_re2_curves): elem (0-based), side
    (1..12), params [n][5], type (one character per record)."""
    path_b = str(path).encode()
    ncurve = C.c_int()
    _lib.call("nrs_re2_info", path_b, None, None, C.byref(ncurve), None)
    n = ncurve.value
    elem, side = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    params = np.zeros(n * 5)
    typ = np.zeros(max(n, 1), dtype=np.uint8)
    if n:
        _lib.call("nrs_re2_curves", path_b, elem.ctypes.data, side.ctypes.data, params.ctypes.data,
                  typ.ctypes.data)
    return {"elem": elem, "side": side, "params": params.reshape(n, 5),
            "type": bytes(typ[:n]).decode("ascii", "replace")}


def mesh_midside(N, vx, vy, vz, curves, x, y, z):
    """Applies midside-node ('m') curved edges to x, y, z in place (nrs_mesh_midside)."""
    vx, vy, vz = (np.ascontiguousarray(a, dtype=np.float64) for a in (vx, vy, vz))
    E = vx.size // 8
    n = len(curves["elem"])
    elem = np.ascontiguousarray(curves["elem"], dtype=np.int32)
    side = np.ascontiguousarray(curves["side"], dtype=np.int32)
    params = np.ascontiguousarray(curves["params"], dtype=np.float64).ravel()
    typ = np.frombuffer(curves["type"].encode("ascii").ljust(max(n, 1)), dtype=np.uint8).copy()
    for a in (x, y, z):
        assert a.dtype == np.float64 and a.flags.c_contiguous and a.size == E * (N + 1) ** 3
    _lib.call("nrs_mesh_midside", N, E, vx.ctypes.data, vy.ctypes.data, vz.ctypes.data, n,
              elem.ctypes.data, side.ctypes.data, params.ctypes.data, typ.ctypes.data,
              x.ctypes.data, y.ctypes.data, z.ctypes.data)


# .re2 boundary codes -> velocity boundary types (src/bdry/bcType.h)
RE2_BC = {"W": 1, "v": 3, "V": 3, "O": 16, "o": 16, "ON": 15, "SYM": 7, "int": 2, "E": 0, "P": 0,
          "": 0}


class Re2Mesh:
    """A mesh read from a .re2 file at order N: coordinates (trilinear in the
    corners, elements with midside-node curved edges triquadratic), global ids
    from the corner topology with periodic ('P') face pairs connected,
    EToB = 1 on faces whose code is in `dirichlet` (for the pressure solve the
    outflow faces), and bc_types for the pressure right-hand side."""

    def __init__(self, path, N, dirichlet=("O", "o", "ON"), scale=1.0):
        d = read_re2(path)
        self.N, self.Nq = N, N + 1
        self.Np = self.Nq ** 3
        self.E = d["nelgt"]
        self.Nlocal = self.E * self.Np
        self.cbc = d["cbc"]
        self.ncurve = d["ncurve"]
        self.peer = d["peer"]
        periodic = self.peer if np.any(self.peer >= 0) else None
        self.x, self.y, self.z, self.gid = mesh_from_vertices(N, d["vx"] * scale, d["vy"] * scale,
                                                              d["vz"] * scale, peer=periodic)
        self.curves = re2_curves(path) if self.ncurve else None
        if self.curves is not None:
            cv = dict(self.curves, params=self.curves["params"] * scale)
            mesh_midside(N, d["vx"] * scale, d["vy"] * scale, d["vz"] * scale, cv, self.x, self.y, self.z)
        codes = np.char.strip(self.cbc.astype(str))
        self.EToB = np.isin(codes, list(dirichlet)).astype(np.int32).ravel()
        unknown = sorted(set(codes.ravel()) - set(RE2_BC))
        if unknown:
            raise ValueError("unsupported boundary codes in %s: %s" % (path, unknown))
        self.bc_types = np.vectorize(RE2_BC.get)(codes).astype(np.int32).ravel()


def coarse_global_ids(Nf, E, ids_fine, Nc):
    """Order-Nc global ids derived from order-Nf ones (host arrays)."""
    ids_fine = np.ascontiguousarray(ids_fine, dtype=np.int64)
    out = np.zeros(E * (Nc + 1) ** 3, dtype=np.int64)
    _lib.call("nrs_coarse_global_ids", Nf, E, ids_fine.ctypes.data, Nc, out.ctypes.data)
    return out


def geometric_factors(N, x, y, z, ggeo, stream=None):
    E = x.numel() // (N + 1) ** 3
    _lib.call("nrs_geometric_factors", N, E, _ptr(x), _ptr(y), _ptr(z), _ptr(ggeo), _stream(stream))


# ------------------------------------------------------------------ kernel seams
def ax(N, ggeo, D, q, Aq, lambda0=None, lambda1=None, poisson=True, lambda_field=False,
       element_list=None, stream=None):
    """Aq = Ax(q) per element (ellipticPartialAxCoeffHex3D)."""
    torch = _torch()
    E = len(element_list) if element_list is not None else q.numel() // (N + 1) ** 3
    if lambda0 is None:
        lambda0 = torch.ones(1, dtype=q.dtype, device=q.device)
    _lib.call("nrs_ax", N, _dtype_code(q), E, _ptr(element_list), _ptr(ggeo), _ptr(D),
              _ptr(lambda0), _ptr(lambda1), int(poisson), int(lambda_field), _ptr(q), _ptr(Aq),
              _stream(stream))


def block_ax(N, ggeo, D, lambda0, lambda1, q, Aq, offset, loffset, E, stream=None):
    _lib.call("nrs_block_ax", N, E, offset, loffset, _ptr(ggeo), _ptr(D), _ptr(lambda0),
              _ptr(lambda1), _ptr(q), _ptr(Aq), _stream(stream))


def build_diagonal(N, ggeo, D, diag, lambda0, lambda1=None, poisson=True, lambda_field=False,
                   stream=None):
    E = diag.numel() // (N + 1) ** 3
    _lib.call("nrs_build_diagonal", N, _dtype_code(diag), E, _ptr(ggeo), _ptr(D), _ptr(lambda0),
              _ptr(lambda1), int(poisson), int(lambda_field), _ptr(diag), _stream(stream))


def mask(mask_ids, v, stream=None):
    _lib.call("nrs_mask", _dtype_code(v), mask_ids.numel(), _ptr(mask_ids), _ptr(v), _stream(stream))


def coarsen(Nf, Nc, R, qf, qc, stream=None):
    E = qf.numel() // (Nf + 1) ** 3
    _lib.call("nrs_coarsen", Nf, Nc, _dtype_code(qf), E, _ptr(R), _ptr(qf), _ptr(qc), _stream(stream))


def prolongate(Nf, Nc, R, qc, qf, stream=None):
    E = qc.numel() // (Nc + 1) ** 3
    _lib.call("nrs_prolongate", Nf, Nc, _dtype_code(qf), E, _ptr(R), _ptr(qc), _ptr(qf),
              _stream(stream))


class Ogs:
    """Gather-scatter handle (ogsSetup / oogs::startFinish)."""

    def __init__(self, global_ids):
        ids = np.ascontiguousarray(global_ids, dtype=np.int64)
        self.N = ids.size
        h = C.c_void_p()
        _lib.call("nrs_ogs_setup", self.N, ids.ctypes.data, C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().nrs_ogs_free(h)
            except Exception:
                pass
            self._h = None

    def gather_scatter(self, v, op="add", nvec=1, stride=None, stream=None):
        stride = self.N if stride is None else stride
        _lib.call("nrs_ogs_gather_scatter", self._h, _ptr(v), nvec, stride, _dtype_code(v),
                  OPS[op], _stream(stream))

    def start(self, v, op="add", nvec=1, stride=None, stream=None):
        """oogs::start: the first half of the split gather-scatter (v is not read)."""
        stride = self.N if stride is None else stride
        _lib.call("nrs_ogs_start", self._h, _ptr(v), nvec, stride, _dtype_code(v), OPS[op], _stream(stream))

    def finish(self, v, op="add", nvec=1, stride=None, stream=None):
        """oogs::finish: the gather-scatter of v as it is now (same arguments as start)."""
        stride = self.N if stride is None else stride
        _lib.call("nrs_ogs_finish", self._h, _ptr(v), nvec, stride, _dtype_code(v), OPS[op], _stream(stream))

    def inv_degree(self, out=None, stream=None):
        torch = _torch()
        if out is None:
            out = torch.empty(self.N, dtype=torch.float64, device="cuda")
        _lib.call("nrs_ogs_inv_degree", self._h, _ptr(out), _stream(stream))
        return out

    def info(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.call("nrs_ogs_info", self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"rows": a.value, "members": b.value, "singletons": c.value}


# velocity boundary types of src/bdry/bcType.h (0: interior face)
BC = {"interior": 0, "wall": 1, "interface": 2, "inlet": 3, "symx": 4, "symy": 5, "symz": 6,
      "sym": 7, "outflow": 16}


REDUCTIONS = {"sum": 0, "max": 1, "min": 2, "norm2": 3, "norm1": 4, "amax": 5,
              "weighted_norm2": 6, "weighted_inner_prod": 7, "inner_prod": 8}


def linalg_reduce(op, x, y=None, w=None, N=None, nfields=1, field_offset=None, stream=None,
                  noalloc=False):
    """linAlg_t reductions (sum, max, min, norm2, norm1, amax, weighted_norm2
    = weightedNorm2Many, weighted_inner_prod = weightedInnerProdMany, inner_prod)
    over N entries of `nfields` fields spaced `field_offset` apart; a float.
    noalloc=True calls the allocation-free entry point (partials in the
    library's fixed device array)."""
    if N is None:
        N = x.numel() // nfields
    off = N if field_offset is None else field_offset
    out = C.c_double()
    _lib.call("nrs_linalg_reduce_noalloc" if noalloc else "nrs_linalg_reduce", REDUCTIONS[op], N,
              nfields, off, _ptr(w), _ptr(x), _ptr(y), C.byref(out), _stream(stream))
    return out.value


def block_update_pcg(inv_degree, Ap, alpha, r, N=None, nfields=1, offset=None, stream=None):
    """ellipticBlockUpdatePCG: r -= alpha Ap in place; returns sum w r^2."""
    if N is None:
        N = inv_degree.numel()
    off = N if offset is None else offset
    out = C.c_double()
    _lib.call("nrs_block_update_pcg", N, nfields, off, _ptr(inv_degree), _ptr(Ap), float(alpha),
              _ptr(r), C.byref(out), _stream(stream))
    return out.value


class PressureRhs:
    """The pressure right-hand side of the TOMBO splitting (tombo::pressureSolve,
    src/navierStokes/tombo.cpp:11-121) on a mesh; bc_types: [E*6] face types
    (BC ids; default: the mesh's own bc_types when it has them (Re2Mesh), else every
    boundary face an inlet, 0 inside)."""

    def __init__(self, mesh, bc_types=None, fusion=True, autotune=True, stress=False,
                 stored_geometry=False):
        """stress: VELOCITY STRESSFORMULATION = TRUE (the stress term of
        pressureStressHex3D taken from the second curl); stored_geometry: read
        the stored geometric factors even on a mesh of trilinear elements."""
        self.mesh = mesh
        self.N, self.E = mesh.N, mesh.E
        self.Nlocal = mesh.E * (mesh.N + 1) ** 3
        if bc_types is None:
            bc_types = getattr(mesh, "bc_types", None)       # a Re2Mesh carries its own
        if bc_types is None:
            bc_types = np.where(np.asarray(mesh.EToB) != 0, BC["inlet"], 0)
        self.bc_types = np.ascontiguousarray(bc_types, dtype=np.int32).reshape(-1)
        assert self.bc_types.size == 6 * mesh.E
        h = C.c_void_p()
        flags = ((1 if fusion else 0) | (2 if autotune else 0) | (4 if stress else 0) |
                 (8 if stored_geometry else 0))
        _lib.call("nrs_tombo_setup", mesh.N, mesh.E, mesh.x.ctypes.data, mesh.y.ctypes.data,
                  mesh.z.ctypes.data, mesh.gid.ctypes.data, self.bc_types.ctypes.data, flags,
                  C.byref(h))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().nrs_tombo_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def fused_pairs(self):
        """Face pairs the curl kernels sum (0: fusion off or not kept by the autotuner)."""
        a, b = C.c_int64(), C.c_int64()
        _lib.call("nrs_tombo_info", self._h, C.byref(a), C.byref(b))
        return a.value

    def _copy(self, which):
        torch = _torch()
        p, n = C.c_void_p(), C.c_int64()
        _lib.call("nrs_tombo_buffer", self._h, which, C.byref(p), C.byref(n))
        out = torch.empty(n.value, dtype=torch.float64, device="cuda")
        _lib.call("nrs_tombo_copy_buffer", self._h, which, _ptr(out), n.value, _stream())
        return out

    @property
    def vgeo(self):
        """[E][10][Np]: rx ry rz sx sy sz tx ty tz JW (a device copy)."""
        return self._copy(0)

    @property
    def inv_lmm(self):
        return self._copy(1)

    @property
    def corner_geometry(self):
        """True when the volume kernels form rx..tz, JW from the element corners
        (trilinear elements, stored_geometry False) instead of streaming vgeo."""
        p, n = C.c_void_p(), C.c_int64()
        _lib.call("nrs_tombo_buffer", self._h, 2, C.byref(p), C.byref(n))
        return n.value > 0

    def curl(self, U, W=None, field_offset=None, stream=None):
        torch = _torch()
        off = self.Nlocal if field_offset is None else field_offset
        if W is None:
            W = torch.empty_like(U)
        _lib.call("nrs_tombo_curl", self._h, off, _ptr(U), _ptr(W), _stream(stream))
        return W

    def compute(self, Ue, U, BF, mue, irho, div, g0idt, visc=True, rhs=None, field_offset=None,
                stream=None):
        torch = _torch()
        off = self.Nlocal if field_offset is None else field_offset
        if rhs is None:
            rhs = torch.empty(self.Nlocal, dtype=torch.float64, device="cuda")
        _lib.call("nrs_tombo_pressure_rhs", self._h, off, _ptr(Ue) if Ue is not None else None,
                  _ptr(U), _ptr(BF), _ptr(mue) if mue is not None else None,
                  _ptr(irho) if irho is not None else None, _ptr(div), float(g0idt), int(visc),
                  _ptr(rhs), _stream(stream))
        return rhs


def fdm_1d(N, lbc, rbc, ll, lm, lr):
    """Host-only 1D FDM pair: (S[nl][nl] (point, eigenvector), lam[nl]), nl = N + 3."""
    import numpy as np
    nl = N + 3
    S = np.zeros(nl * nl)
    lam = np.zeros(nl)
    _lib.call("nrs_fdm_1d", N, lbc, rbc, float(ll), float(lm), float(lr), S.ctypes.data,
              lam.ctypes.data)
    return S.reshape(nl, nl), lam


def fdm_element_lengths(mesh):
    """Host-only element lengths of a mesh: (left, middle, right), each [3][E]."""
    import numpy as np
    out = np.zeros(9 * mesh.E)
    _lib.call("nrs_fdm_element_lengths", mesh.N, mesh.E, mesh.x.ctypes.data, mesh.y.ctypes.data,
              mesh.z.ctypes.data, mesh.gid.ctypes.data, out.ctypes.data)
    out = out.reshape(3, 3, mesh.E)
    return out[0], out[1], out[2]


def default_opts():
    o = _lib.EllipticOpts()
    _lib.load().nrs_elliptic_default_opts(C.byref(o))
    return o


COARSE = {"pcg": 0, "smooth": 1, "smooth_jacobi": 2, "amg": 3, "user": 4}
INITIAL_GUESS = {"previous": 0, "projection": 1, "projection-aconj": 2}


def make_desc(mesh, lambda0=1.0, lambda1=0.0, poisson=True, nfields=1, lambda0_field=None,
              lambda1_field=None):
    """nrs_elliptic_desc of a mesh (host arrays stay owned by the mesh);
    lambda0_field / lambda1_field: per-node FP64 device tensors [E*Np] (ELLIPTIC
    COEFF FIELD), referenced, not copied: the caller keeps them alive."""
    return _lib.EllipticDesc(mesh.N, mesh.E, mesh.x.ctypes.data, mesh.y.ctypes.data,
                             mesh.z.ctypes.data, mesh.gid.ctypes.data, mesh.EToB.ctypes.data,
                             int(poisson), float(lambda0), float(lambda1), int(nfields),
                             _field_ptr(lambda0_field), _field_ptr(lambda1_field))


def _field_ptr(t):
    if t is None:
        return None
    assert t.is_cuda and t.dtype == _torch().float64 and t.is_contiguous(), "coefficient field: FP64 device"
    return t.data_ptr()


def parse_mg_schedule(schedule, N, default_degree=3, opts=None):
    """MULTIGRID SCHEDULE (pMGSchedule), e.g. "p=7+degree=2, p=3+degree=2, p=1,
    p=3+degree=2, p=7+degree=2", into opts (a new default one if None): levels
    and the per-level degrees before restriction / after prolongation."""
    o = default_opts() if opts is None else opts
    _lib.call("nrs_parse_mg_schedule", schedule.encode(), int(N), int(default_degree), C.byref(o))
    return o


def make_opts(solver="pcg", precon="jacobi", tol=1e-8, relative=True, max_iter=500, levels=None,
              smoother="chebyshev+jacobi", cheby_degree=2, cheby_min=0.1, cheby_max=1.1,
              coarse_tol=1e-2, coarse_max_iter=200, seed=0, coarse="pcg",
              initial_guess="previous", proj_vectors=8, proj_start=5, schedule=None, order=None,
              **extra):
    """nrs_elliptic_opts from keyword options (extra: raw field names).
    initial_guess: "previous", "projection" or "projection-aconj" (INITIAL GUESS);
    proj_vectors / proj_start: RESIDUAL PROJECTION VECTORS / START;
    schedule: a MULTIGRID SCHEDULE string (needs order = the mesh order; the
    entries without a degree take cheby_degree), in place of levels."""
    o = default_opts()
    o.solver = SOLVERS[solver]
    o.preconditioner = PRECONS[precon]
    o.tol = tol
    o.relative = int(relative)
    o.maxIter = max_iter
    if levels:
        o.nLevels = len(levels)
        for i, l in enumerate(levels):
            o.levels[i] = l
    o.smoother = SMOOTHERS[smoother]
    o.chebyDegree = cheby_degree
    o.chebyMinFactor = cheby_min
    o.chebyMaxFactor = cheby_max
    o.coarseTol = coarse_tol
    o.coarseMaxIter = coarse_max_iter
    o.coarseSolver = COARSE[coarse]
    o.seed = seed
    o.initialGuess = INITIAL_GUESS[initial_guess]
    o.projVectors = proj_vectors
    o.projStart = proj_start
    if schedule is not None:
        parse_mg_schedule(schedule, order, cheby_degree, o)
    for k, v in extra.items():
        setattr(o, k, v)
    return o


TUNE_KERNELS = {"ax": 0, "fdm": 1, "op": 2, "tombo": 3}


def _tune_type(kernel, dtype):
    if kernel == "fdm":
        return FDM_PRECISIONS[dtype]
    if kernel == "tombo":
        return {"curl": 0, "rhs": 1}[dtype]
    return DTYPES[dtype]


def tuned_variant(kernel, N, dtype, nelements):
    """Recorded launch variant for (kernel, N, type, E), 0 if none (nrs_tuned_variant).
    type: a DTYPES name for "ax" and "op", an FDM_PRECISIONS name for "fdm", "curl" or
    "rhs" for "tombo"."""
    t = _tune_type(kernel, dtype)
    v = C.c_int()
    _lib.call("nrs_tuned_variant", TUNE_KERNELS[kernel], int(N), t, int(nelements), C.byref(v))
    return v.value


def set_variant(kernel, N, dtype, nelements, variant):
    """Force a launch variant (0 clears) for later launches with this key."""
    t = _tune_type(kernel, dtype)
    _lib.call("nrs_set_variant", TUNE_KERNELS[kernel], int(N), t, int(nelements), int(variant))


class Batch:
    """Independent solves advanced together by one captured graph (nrs_batch_*).

    problems: list of (mesh, dict(lambda0=, lambda1=, poisson=, **solver options))."""

    def __init__(self, problems, iters_per_launch=1):
        n = len(problems)
        self.meshes = [m for m, _ in problems]
        descs = (_lib.EllipticDesc * n)()
        opts = (_lib.EllipticOpts * n)()
        for i, (m, kw) in enumerate(problems):
            kw = dict(kw)
            lam = {k: kw.pop(k) for k in ("lambda0", "lambda1", "poisson") if k in kw}
            descs[i] = make_desc(m, **lam)
            opts[i] = make_opts(**kw)
        h = C.c_void_p()
        _lib.call("nrs_batch_setup", n, descs, opts, int(iters_per_launch), C.byref(h))
        self._h = h
        self.n = n

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().nrs_batch_free(h)
            except Exception:
                pass
            self._h = None

    def info(self):
        n, nodes, launches = C.c_int(), C.c_int64(), C.c_int()
        _lib.call("nrs_batch_info", self._h, C.byref(n), C.byref(nodes), C.byref(launches))
        return {"problems": n.value, "graph_nodes": nodes.value, "last_launches": launches.value}

    def set_mode(self, mode):
        """"one_graph" (all problems in one graph) or "per_problem" (a graph per
        problem, active problems only)."""
        _lib.call("nrs_batch_set_mode", self._h, {"one_graph": 0, "per_problem": 1}[mode])

    def problem(self, p):
        """Problem p's solver handle (borrowed: valid while the batch lives)."""
        h = C.c_void_p()
        _lib.call("nrs_batch_problem", self._h, int(p), C.byref(h))
        return Elliptic._borrow(h, self.meshes[p], self)

    def solve(self, rhs, x, stream=None):
        """In place per problem (as Elliptic.solve); returns a list of stats dicts."""
        n = self.n
        rp = (C.c_void_p * n)(*[_ptr(t) for t in rhs])
        xp = (C.c_void_p * n)(*[_ptr(t) for t in x])
        st = (_lib.SolveStats * n)()
        _lib.call("nrs_batch_solve", self._h, rp, xp, st, _stream(stream))
        return [{"iterations": s.iterations, "converged": bool(s.converged), "res0": s.res0,
                 "res": s.res, "seconds": s.seconds} for s in st]


class Elliptic:
    """elliptic_t for one scalar field (ellipticSolveSetup / ellipticSolve)."""

    def __init__(self, mesh, lambda0=1.0, lambda1=0.0, poisson=True, nfields=1, lambda0_field=None,
                 lambda1_field=None, **kw):
        """nfields=3: block solve of three fields (vectors [3][Nlocal]), Jacobi or
        no preconditioner; Nlocal stays the per-field length.  lambda0_field /
        lambda1_field: per-node coefficients (FP64 device tensors [E*Np], kept
        by reference: changing their values changes the operator)."""
        self.mesh = mesh
        self.nfields = nfields
        self._fields = (lambda0_field, lambda1_field)
        d = make_desc(mesh, lambda0, lambda1, poisson, nfields, lambda0_field, lambda1_field)
        if kw.get("schedule") is not None:
            kw.setdefault("order", mesh.N)
        o = make_opts(**kw)
        self.opts = o
        h = C.c_void_p()
        _lib.call("nrs_elliptic_setup", C.byref(d), C.byref(o), C.byref(h))
        self._h = h
        n, m, vol = C.c_int64(), C.c_int64(), C.c_double()
        _lib.call("nrs_elliptic_info", h, C.byref(n), C.byref(m), C.byref(vol))
        self.Nlocal, self.Nmasked, self.volume = n.value, m.value, vol.value

    @classmethod
    def _borrow(cls, h, mesh, owner):
        """Wrap a handle owned elsewhere (kept alive through `owner`, not freed here)."""
        self = cls.__new__(cls)
        self.mesh, self._h, self._owner = mesh, h, owner
        n, m, vol = C.c_int64(), C.c_int64(), C.c_double()
        _lib.call("nrs_elliptic_info", h, C.byref(n), C.byref(m), C.byref(vol))
        self.Nlocal, self.Nmasked, self.volume = n.value, m.value, vol.value
        return self

    def __del__(self):
        h = getattr(self, "_h", None)
        if getattr(self, "_owner", None) is not None:
            return
        if h is not None and h.value:
            try:
                _lib.load().nrs_elliptic_free(h)
            except Exception:
                pass
            self._h = None

    def buffer(self, which):
        """Device tensor view of an internal buffer (copy)."""
        torch = _torch()
        p, n = C.c_void_p(), C.c_int64()
        _lib.call("nrs_elliptic_buffer", self._h, which, C.byref(p), C.byref(n))
        dt = torch.int32 if which == 3 else torch.float64
        out = torch.empty(n.value, dtype=dt, device="cuda")
        _lib.call("nrs_elliptic_copy_buffer", self._h, which, _ptr(out), n.value, _stream())
        return out

    @property
    def ggeo(self):
        return self.buffer(0)

    @property
    def D(self):
        return self.buffer(1)

    @property
    def inv_degree(self):
        return self.buffer(2)

    @property
    def mask_ids(self):
        return self.buffer(3)

    def set_coefficients(self, lambda0_field=None, lambda1_field=None, stream=None):
        """Point the operator at new coefficient fields (None, None: the
        constants) and rebuild the Jacobi diagonal."""
        _lib.call("nrs_elliptic_set_coefficients", self._h, _field_ptr(lambda0_field),
                  _field_ptr(lambda1_field), _stream(stream))
        self._fields = (lambda0_field, lambda1_field)

    def update_jacobi(self, stream=None):
        """ellipticUpdateJacobi: the Jacobi diagonal from the coefficients' current values."""
        _lib.call("nrs_elliptic_update_jacobi", self._h, _stream(stream))

    def operator(self, q, Aq, masked=True, stream=None):
        _lib.call("nrs_elliptic_operator", self._h, _ptr(q), _ptr(Aq), int(masked), _stream(stream))

    @property
    def fused_pairs(self):
        """Face pairs summed inside the Ax kernel (operator face fusion)."""
        n = C.c_int64()
        _lib.call("nrs_elliptic_fusion_info", self._h, C.byref(n))
        return n.value

    def preconditioner(self, r, z, stream=None):
        rc = _lib.load().nrs_elliptic_preconditioner(self._h, _ptr(r), _ptr(z), _stream(stream))
        self._raise_user_error()
        _lib.check(rc)

    def mg_info(self):
        """Multigrid levels: [{order, Nlocal, rho, degree}], coarse PCG iterations."""
        n = C.c_int()
        orders = (C.c_int * 8)(); nl = (C.c_int64 * 8)(); rho = (C.c_double * 8)()
        deg = (C.c_int * 8)(); cit = C.c_int()
        _lib.call("nrs_elliptic_mg_info", self._h, 8, C.byref(n), orders, nl, rho, deg, C.byref(cit))
        lv = [{"order": orders[i], "Nlocal": nl[i], "rho": rho[i], "degree": deg[i]}
              for i in range(min(n.value, 8))]
        return lv, cit.value

    def mg_smoother(self, level, x, out, stream=None):
        """out = S x on a non-coarse multigrid level (tensors of the MG precision)."""
        _lib.call("nrs_elliptic_mg_smoother", self._h, int(level), _ptr(x), _ptr(out),
                  _stream(stream))

    MG_LEVEL_OPS = {"ax": 0, "residual": 1, "smooth": 2, "coarsen": 3, "prolongate": 4}

    def mg_level(self, level, op, inp, out, x=None, x_is_zero=True, stream=None):
        """One multigridLevel operation on `level` (tensors of the MG precision):
        "ax" out = A inp; "residual" out = inp - A x; "smooth" smooths out for
        the rhs inp; "coarsen" (level >= 1) out = R (inp .* 1/multiplicity of
        level-1, inp overwritten); "prolongate" (level >= 1) out += P inp."""
        _lib.call("nrs_elliptic_mg_level", self._h, int(level), self.MG_LEVEL_OPS[op], _ptr(inp),
                  _ptr(x), _ptr(out), int(x_is_zero), _stream(stream))

    def profile(self, on=True):
        """Enable/disable per-launch device timing of the operator's kernels."""
        _lib.call("nrs_elliptic_profile_enable", self._h, int(on))

    def profile_read(self):
        """{phase: (total_ms, launches)} since profiling was enabled."""
        n = len(_lib.PHASES)
        ms = (C.c_double * n)()
        cnt = (C.c_int64 * n)()
        _lib.call("nrs_elliptic_profile_read", self._h, ms, cnt)
        return {p: (ms[i], cnt[i]) for i, p in enumerate(_lib.PHASES)}

    def set_user_preconditioner(self, fn, tensors=True):
        """The PRECONDITIONER = USER hook (elliptic_t::userPreconditioner,
        elliptic.h:178).  fn(r, z, stream) must enqueue z = M^-1 r on `stream`
        (a torch.cuda.ExternalStream); r and z are float64 device tensors of
        Nfields * Nlocal entries viewing the solver's own vectors (tensors=True),
        or their device addresses as ints (tensors=False).  An exception raised
        by fn is re-raised by the solve that called it.  fn=None clears the hook."""
        if fn is None:
            self._user_cb = None
            _lib.call("nrs_elliptic_set_user_preconditioner", self._h, _lib.USER_PRECON_FN(), None)
            return
        torch = _torch()
        n = getattr(self, "nfields", 1) * self.Nlocal
        self._user_err = None

        def cb(user, r, z, stream):
            if self._user_err is not None:
                return
            try:
                # a NULL stream (ctypes gives None) is the device's default stream
                st = torch.cuda.ExternalStream(stream) if stream else torch.cuda.default_stream()
                if tensors:
                    fn(_device_view(r, n), _device_view(z, n), st)
                else:
                    fn(r, z, st)
            except BaseException as e:       # raised again after the C call returns
                self._user_err = e

        self._user_cb = _lib.USER_PRECON_FN(cb)
        _lib.call("nrs_elliptic_set_user_preconditioner", self._h, self._user_cb, None)

    def set_user_coarse_solver(self, fn, tensors=True):
        """The coarse-solve hook of coarse="user" (coarseLevel_t::solvePtr,
        MG/MGSolver.hpp:99): fn(rhs, x, stream) must enqueue the coarsest level's
        solve on `stream`, writing all of x; rhs and x are device tensors of the
        coarsest level in the multigrid precision (tensors=True) or their
        addresses (tensors=False).  Exceptions are re-raised by the solve or
        preconditioner call that ran it.  fn=None clears the hook."""
        if fn is None:
            self._coarse_cb = None
            _lib.call("nrs_elliptic_set_user_coarse_solver", self._h, _lib.USER_COARSE_FN(), None)
            return
        torch = _torch()
        n = self.mg_info()[0][-1]["Nlocal"]
        ts = "<f4" if self.opts.mgPrecision == 0 else "<f8"
        self._user_err = None

        def cb(user, rhs, x, stream):
            if self._user_err is not None:
                return
            try:
                st = torch.cuda.ExternalStream(stream) if stream else torch.cuda.default_stream()
                if tensors:
                    fn(_device_view(rhs, n, ts), _device_view(x, n, ts), st)
                else:
                    fn(rhs, x, st)
            except BaseException as e:
                self._user_err = e

        self._coarse_cb = _lib.USER_COARSE_FN(cb)
        _lib.call("nrs_elliptic_set_user_coarse_solver", self._h, self._coarse_cb, None)

    def _raise_user_error(self):
        e = getattr(self, "_user_err", None)
        if e is not None:
            self._user_err = None
            raise RuntimeError("user preconditioner failed") from e

    def krylov(self, r, x, tol, max_iter, rdotr, stream=None):
        """pcg / pgmres alone (PCG.cpp:85-203, PGMRES.cpp:111-340) on the masked,
        assembled residual r (overwritten); x += the increment.  tol absolute on
        the normalised norm, rdotr = ||r||_w / sqrt(volume) on entry.
        Returns (iterations, final normalised norm)."""
        res = C.c_double(rdotr)
        it = C.c_int()
        rc = _lib.load().nrs_elliptic_krylov(self._h, _ptr(r), _ptr(x), float(tol), int(max_iter),
                                             C.byref(res), C.byref(it), _stream(stream))
        self._raise_user_error()
        _lib.check(rc)
        return it.value, res.value

    def solve(self, rhs, x, stream=None):
        """In place: x <- solution, rhs <- final residual.  Returns a stats dict."""
        st = _lib.SolveStats()
        rc = _lib.load().nrs_elliptic_solve(self._h, _ptr(rhs), _ptr(x), C.byref(st), _stream(stream))
        self._raise_user_error()
        _lib.check(rc)
        return {"iterations": st.iterations, "converged": bool(st.converged), "res0": st.res0,
                "res": st.res, "seconds": st.seconds, "precon_seconds": st.precon_seconds,
                "res00": st.res00, "proj_vectors": st.projVectors}
