"""ctypes binding of libnekrs_hip.so (the C ABI in include/nrs_hip.h).

The shared library is built in-tree by ``nekrs_amd/build.py``.  There is no
fallback: if the library is missing or fails to load, importing this module
raises, so nothing can silently run on the CPU instead.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnekrs_hip.so")

NRS_OK, NRS_ERR_INVALID, NRS_ERR_HIP, NRS_ERR_NAN, NRS_ERR_UNSUPPORTED, NRS_ERR_NO_DEVICE = (
    0, -1, -2, -3, -4, -5)
F64, F32, I32 = 0, 1, 2
ADD, MUL, MIN, MAX = 0, 1, 2, 3


class NrsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("nrs error %d: %s" % (code, msg))
        self.code = code


class EllipticDesc(C.Structure):
    _fields_ = [("N", C.c_int), ("Nelements", C.c_int),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p),
                ("globalIds", C.c_void_p), ("EToB", C.c_void_p),
                ("poisson", C.c_int), ("lambda0", C.c_double), ("lambda1", C.c_double),
                ("Nfields", C.c_int), ("lambda0Field", C.c_void_p), ("lambda1Field", C.c_void_p)]


class EllipticOpts(C.Structure):
    _fields_ = [("solver", C.c_int), ("preconditioner", C.c_int), ("tol", C.c_double),
                ("relative", C.c_int), ("maxIter", C.c_int), ("nLevels", C.c_int),
                ("levels", C.c_int * 8), ("smoother", C.c_int), ("chebyDegree", C.c_int),
                ("chebyMinFactor", C.c_double), ("chebyMaxFactor", C.c_double),
                ("coarseSolver", C.c_int), ("coarseTol", C.c_double), ("coarseMaxIter", C.c_int),
                ("mgPrecision", C.c_int), ("fdmPrecision", C.c_int), ("seed", C.c_ulonglong),
                ("useGraphs", C.c_int), ("autotune", C.c_int), ("gmresRestart", C.c_int),
                ("initialGuess", C.c_int), ("projVectors", C.c_int), ("projStart", C.c_int),
                ("operatorFusion", C.c_int), ("preDegree", C.c_int * 8), ("postDegree", C.c_int * 8),
                ("precoCoeffField", C.c_int), ("mgGeometry", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("res0", C.c_double),
                ("res", C.c_double), ("seconds", C.c_double), ("precon_seconds", C.c_double),
                ("res00", C.c_double), ("projVectors", C.c_int)]


# exported symbols and their signatures (all return int status unless noted)
_vp, _i, _i64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_double
# nrs_user_precon_fn: fn(user, r, z, stream), no return value
USER_PRECON_FN = C.CFUNCTYPE(None, _vp, _vp, _vp, _vp)
# nrs_user_coarse_fn: fn(user, rhs, x, stream)
USER_COARSE_FN = C.CFUNCTYPE(None, _vp, _vp, _vp, _vp)
SIGNATURES = {
    "nrs_last_error": ([], C.c_char_p),
    "nrs_version": ([], C.c_char_p),
    "nrs_init": ([_i], _i),
    "nrs_box_mesh": ([_i, _i, _i, _i, _i, _d, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_gll": ([_i, _vp, _vp, _vp], _i),
    "nrs_geometric_factors": ([_i, _i, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_ax": ([_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp, _vp], _i),
    "nrs_block_ax": ([_i, _i, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_build_diagonal": ([_i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp], _i),
    "nrs_mask": ([_i, _i64, _vp, _vp, _vp], _i),
    "nrs_coarsen": ([_i, _i, _i, _i, _vp, _vp, _vp, _vp], _i),
    "nrs_prolongate": ([_i, _i, _i, _i, _vp, _vp, _vp, _vp], _i),
    "nrs_ogs_setup": ([_i64, _vp, C.POINTER(_vp)], _i),
    "nrs_ogs_free": ([_vp], _i),
    "nrs_ogs_gather_scatter": ([_vp, _vp, _i, _i64, _i, _i, _vp], _i),
    "nrs_ogs_start": ([_vp, _vp, _i, _i64, _i, _i, _vp], _i),
    "nrs_ogs_finish": ([_vp, _vp, _i, _i64, _i, _i, _vp], _i),
    "nrs_ogs_inv_degree": ([_vp, _vp, _vp], _i),
    "nrs_ogs_info": ([_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)], _i),
    "nrs_linalg_reduce": ([_i, _i64, _i, _i64, _vp, _vp, _vp, C.POINTER(_d), _vp], _i),
    "nrs_linalg_reduce_noalloc": ([_i, _i64, _i, _i64, _vp, _vp, _vp, C.POINTER(_d), _vp], _i),
    "nrs_block_update_pcg": ([_i64, _i, _i64, _vp, _vp, _d, _vp, C.POINTER(_d), _vp], _i),
    "nrs_elliptic_default_opts": ([C.POINTER(EllipticOpts)], None),
    "nrs_elliptic_setup": ([C.POINTER(EllipticDesc), C.POINTER(EllipticOpts), C.POINTER(_vp)], _i),
    "nrs_elliptic_free": ([_vp], _i),
    "nrs_elliptic_info": ([_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_d)], _i),
    "nrs_elliptic_fusion_info": ([_vp, C.POINTER(_i64)], _i),
    "nrs_elliptic_buffer": ([_vp, _i, C.POINTER(_vp), C.POINTER(_i64)], _i),
    "nrs_elliptic_copy_buffer": ([_vp, _i, _vp, _i64, _vp], _i),
    "nrs_elliptic_set_coefficients": ([_vp, _vp, _vp, _vp], _i),
    "nrs_elliptic_update_jacobi": ([_vp, _vp], _i),
    "nrs_elliptic_operator": ([_vp, _vp, _vp, _i, _vp], _i),
    "nrs_elliptic_preconditioner": ([_vp, _vp, _vp, _vp], _i),
    "nrs_elliptic_solve": ([_vp, _vp, _vp, C.POINTER(SolveStats), _vp], _i),
    "nrs_elliptic_set_user_preconditioner": ([_vp, USER_PRECON_FN, _vp], _i),
    "nrs_elliptic_mg_level": ([_vp, _i, _i, _vp, _vp, _vp, _i, _vp], _i),
    "nrs_parse_mg_schedule": ([C.c_char_p, _i, _i, C.POINTER(EllipticOpts)], _i),
    "nrs_elliptic_set_user_coarse_solver": ([_vp, USER_COARSE_FN, _vp], _i),
    "nrs_re2_curves": ([C.c_char_p, _vp, _vp, _vp, _vp], _i),
    "nrs_mesh_midside": ([_i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_re2_periodic": ([C.c_char_p, _vp], _i),
    "nrs_mesh_from_vertices_periodic": ([_i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_elliptic_krylov": ([_vp, _vp, _vp, _d, _i, C.POINTER(_d), C.POINTER(_i), _vp], _i),
    "nrs_elliptic_mg_info": ([_vp, _i, C.POINTER(_i), _vp, _vp, _vp, _vp, C.POINTER(_i)], _i),
    "nrs_elliptic_mg_smoother": ([_vp, _i, _vp, _vp, _vp], _i),
    "nrs_fdm_1d": ([_i, _i, _i, _d, _d, _d, _vp, _vp], _i),
    "nrs_fdm_element_lengths": ([_i, _i, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_coarse_global_ids": ([_i, _i, _vp, _i, _vp], _i),
    "nrs_batch_setup": ([_i, _vp, _vp, _i, C.POINTER(_vp)], _i),
    "nrs_batch_solve": ([_vp, _vp, _vp, _vp, _vp], _i),
    "nrs_batch_info": ([_vp, C.POINTER(_i), C.POINTER(_i64), C.POINTER(_i)], _i),
    "nrs_batch_free": ([_vp], _i),
    "nrs_batch_problem": ([_vp, _i, C.POINTER(_vp)], _i),
    "nrs_batch_set_mode": ([_vp, _i], _i),
    "nrs_tuned_variant": ([_i, _i, _i, _i64, C.POINTER(_i)], _i),
    "nrs_set_variant": ([_i, _i, _i, _i64, _i], _i),
    "nrs_elliptic_profile_enable": ([_vp, _i], _i),
    "nrs_elliptic_profile_read": ([_vp, C.POINTER(_d), C.POINTER(_i64)], _i),
    "nrs_tombo_setup": ([_i, _i, _vp, _vp, _vp, _vp, _vp, _i, C.POINTER(_vp)], _i),
    "nrs_tombo_free": ([_vp], _i),
    "nrs_tombo_pressure_rhs": ([_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _d, _i, _vp, _vp], _i),
    "nrs_tombo_curl": ([_vp, _i64, _vp, _vp, _vp], _i),
    "nrs_tombo_info": ([_vp, C.POINTER(_i64), C.POINTER(_i64)], _i),
    "nrs_tombo_buffer": ([_vp, _i, C.POINTER(_vp), C.POINTER(_i64)], _i),
    "nrs_tombo_copy_buffer": ([_vp, _i, _vp, _i64, _vp], _i),
    "nrs_re2_info": ([C.c_char_p, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)], _i),
    "nrs_re2_read": ([C.c_char_p, _vp, _vp, _vp, _vp, _vp], _i),
    "nrs_mesh_from_vertices": ([_i, _i, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp], _i),
}
PHASES = ("ax", "gs")

_lib = None


def load():
    """Load libnekrs_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libnekrs_hip.so is not built: run `python nekrs_amd/build.py` "
                          "(or __graft_entry__.build()) first; there is no CPU fallback")
    # One HIP runtime per process: torch ships its own libamdhip64.so.7, and
    # whichever copy is loaded first serves both.  Loading torch first makes
    # that torch's, so its streams and allocations are the library's too
    # (loading the system copy first leaves torch without devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc):
    if rc != NRS_OK:
        raise NrsError(rc, load().nrs_last_error().decode())
    return rc


def call(name, *args):
    return check(getattr(load(), name)(*args))
