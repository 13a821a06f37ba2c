"""ctypes binding of liborbit_hip.so (the C ABI declared in include/orbit_hip.h).

The library is loaded from this package directory only; there is no CPU
fallback.  Any call made without it raises ``OrbitLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborbit_hip.so")

ORBIT_OK = 0
ORBIT_WARN_NOT_CONVERGED = 1
ORBIT_ERR_INVALID = -1
ORBIT_ERR_HIP = -2
ORBIT_ERR_STRUCTURE = -3
ORBIT_ERR_NOMEM = -4
ORBIT_PREC_FP64 = 0
ORBIT_PREC_MIXED = 1
ORBIT_SWEEP_ACC_MIXED = 2

# Every symbol include/orbit_hip.h declares.
EXPORTS = [
    "orbit_last_error", "orbit_opts_default", "orbit_ctx_create", "orbit_ctx_destroy",
    "orbit_ctx_sync", "orbit_ctx_stream", "orbit_opset_create_csr", "orbit_opset_create_shared",
    "orbit_opset_destroy", "orbit_opset_stats_get", "orbit_opset_update_diag", "orbit_apply_P",
    "orbit_solve_forward", "orbit_solve_adjoint", "orbit_solve_forward_dev",
    "orbit_solve_adjoint_dev", "orbit_mass_balance", "orbit_fct_source", "orbit_fct_source_dev",
    "orbit_fct_linop_create", "orbit_fct_linop_apply", "orbit_fct_linop_apply_dev",
    "orbit_fct_linop_csr", "orbit_fct_linop_destroy", "orbit_bench_halfsweep",
    "orbit_fct_grid_create", "orbit_fct_grid_destroy", "orbit_fct_source_batch_dev",
    "orbit_fct_linop_create_dev", "orbit_opset_check", "orbit_spmat_create",
    "orbit_spmat_apply_dev", "orbit_spmat_destroy", "orbit_opset_autotune",
    "orbit_opset_get_shapes", "orbit_opset_set_shapes", "orbit_transport_assemble",
    "orbit_opset_create_grids", "orbit_lut_create", "orbit_lut_query_dev", "orbit_lut_destroy",
    "orbit_species_diagonals", "orbit_lateral_inflow", "orbit_opset_set_relaxation", "orbit_forward_fct_dev",
    "orbit_solve_forward_dev_noalloc", "orbit_timing_start", "orbit_timing_stop",
    "orbit_forward_fct_passes_dev", "orbit_lut_partitioning_dev", "orbit_iso_cross_blocks_dev", "orbit_opset_set_mfma",
    "orbit_solve_group_dev", "orbit_opset_set_patch", "orbit_opset_set_handoff",
]
N_SWEEP_PRECS = 3    # rows of a shape table: FP64, accurate mixed, FP32 inner sweeps
N_SLOT_BUCKETS = 5   # columns: 1, 2-3, 4-7, 8-15, 16+ slots


class OrbitLibraryMissing(RuntimeError):
    pass


class OrbitError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"orbit error {code}: {msg}")
        self.code = code


class Opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxiter", C.c_int), ("restart", C.c_int),
                ("precision", C.c_int), ("inner_sweeps", C.c_int), ("acc_sweeps", C.c_int),
                ("use_graph", C.c_int), ("verbose", C.c_int), ("max_rounds", C.c_int),
                ("rhs_is_s", C.c_int), ("warm_orbit", C.c_int), ("check_sweeps", C.c_int)]


class Info(C.Structure):
    _fields_ = [("gmres_iters", C.c_int), ("rounds", C.c_int), ("converged", C.c_int),
                ("zero_source", C.c_int), ("residual", C.c_double), ("periodicity", C.c_double),
                ("neg_frac", C.c_double), ("s_norm", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SweepTiming(C.Structure):
    _fields_ = [("precision", C.c_int), ("slots_per_wave", C.c_int), ("mfma", C.c_int), ("own_zero", C.c_int),
                ("nbr_zero", C.c_int), ("launches", C.c_longlong), ("total_ms", C.c_double),
                ("total_bytes", C.c_double), ("first_sweep", C.c_int)]


class OpsetStats(C.Structure):
    _fields_ = [("nz", C.c_int), ("ny", C.c_int), ("nx", C.c_int), ("n_bins", C.c_int),
                ("n_tracers", C.c_int), ("n_groups", C.c_int),
                ("default_inner_sweeps", C.c_int), ("default_acc_sweeps", C.c_int),
                ("dtau", C.c_double), ("device_bytes", C.c_double), ("sweep_contraction", C.c_double),
                ("periodic_x", C.c_int), ("relaxation", C.c_double), ("mfma", C.c_int),
                ("default_check_sweeps", C.c_int), ("patch_shape", C.c_int)]


class MbStats(C.Structure):
    _fields_ = [("interior_median", C.c_double), ("interior_p90", C.c_double),
                ("interior_p99", C.c_double), ("interior_max", C.c_double),
                ("interior_frac_above_0_01", C.c_double), ("interior_frac_above_0_1", C.c_double),
                ("top_median", C.c_double), ("top_p90", C.c_double),
                ("layer_p90", C.c_double * 32), ("layer_mean_signed", C.c_double * 32)]


class GridMetrics(C.Structure):
    _fields_ = [("nz", C.c_int), ("ny", C.c_int), ("nx", C.c_int), ("split_fluxes", C.c_int),
                ("concentration_state", C.c_int),
                ("u_plus", C.c_void_p), ("u_minus", C.c_void_p), ("v_plus", C.c_void_p),
                ("v_minus", C.c_void_p), ("dx", C.c_void_p), ("dy", C.c_double),
                ("dP", C.c_void_p), ("Dz", C.c_void_p)]


class SweepShape(C.Structure):
    _fields_ = [("slots_per_wave", C.c_int), ("threads", C.c_int), ("xcd_map", C.c_int)]


ShapeTable = SweepShape * (N_SWEEP_PRECS * N_SLOT_BUCKETS)

_lib = None
_lock = threading.Lock()
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PP = C.POINTER(C.c_void_p)


def signatures():
    """{symbol: (restype, argtypes)} of every entry point include/orbit_hip.h declares."""
    return {
        "orbit_last_error": (C.c_char_p, []),
        "orbit_opts_default": (None, [C.POINTER(Opts)]),
        "orbit_ctx_create": (_I, [_I, C.POINTER(_P)]),
        "orbit_ctx_destroy": (_I, [_P]),
        "orbit_ctx_sync": (_I, [_P]),
        "orbit_ctx_stream": (_P, [_P]),
        "orbit_opset_create_csr": (_I, [_P, _I, _I, _I, _I, _I, _D, _PP, _PP, _PP, C.POINTER(_P)]),
        "orbit_opset_create_shared": (_I, [_P, _I, _I, _I, _I, _I, _D, _PP, _PP, _PP, _P, C.POINTER(_P)]),
        "orbit_opset_destroy": (_I, [_P]),
        "orbit_opset_stats_get": (_I, [_P, C.POINTER(OpsetStats)]),
        "orbit_opset_update_diag": (_I, [_P, _P, _P]),
        "orbit_apply_P": (_I, [_P, _P, _I, _I, _I, _P, _P]),
        "orbit_solve_forward": (_I, [_P, _P, _P, _I, _P, _P, C.POINTER(Opts), _P, _P]),
        "orbit_solve_adjoint": (_I, [_P, _P, _P, _I, _P, _P, C.POINTER(Opts), _P, _P]),
        "orbit_solve_forward_dev": (_I, [_P, _P, _P, _I, _P, _P, C.POINTER(Opts), _P, _P]),
        "orbit_solve_forward_dev_noalloc": (_I, [_P, _P, _P, _I, _P, _P, C.POINTER(Opts), _P, _P]),
        "orbit_iso_cross_blocks_dev": (_I, [_P, _I, _I, _I, _I, C.c_longlong, _P, _P, _P, _P, _P, _P, _P, C.c_uint,
                                           _P]),
        "orbit_lut_partitioning_dev": (_I, [_P, _P, _I, _I, _I, C.c_longlong, _P, C.c_longlong, _P, C.c_longlong,
                                            _I, C.c_double, _P, _P]),
        "orbit_opset_set_mfma": (_I, [_P, _I]),
        "orbit_opset_set_patch": (_I, [_P, _I]),
        "orbit_opset_set_handoff": (_I, [_P, _I]),
        "orbit_solve_group_dev": (_I, [_P, _I, _P, _I, _P, _P, _P, _I, C.POINTER(Opts), _P, _P]),
        "orbit_timing_start": (_I, [_P, _I]),
        "orbit_timing_stop": (_I, [_P, _P, _I, _P, _P]),
        "orbit_solve_adjoint_dev": (_I, [_P, _P, _P, _I, _P, _P, C.POINTER(Opts), _P, _P]),
        "orbit_mass_balance": (_I, [_P, _P, _P, C.POINTER(MbStats)]),
        "orbit_fct_source": (_I, [_P, C.POINTER(GridMetrics), _P, _D, _I, _P, _P]),
        "orbit_fct_source_dev": (_I, [_P, C.POINTER(GridMetrics), _P, _D, _I, _P, _P]),
        "orbit_fct_linop_create": (_I, [_P, C.POINTER(GridMetrics), _P, _D, C.POINTER(_P)]),
        "orbit_fct_linop_apply": (_I, [_P, _P, _I, _P, _P]),
        "orbit_fct_linop_apply_dev": (_I, [_P, _P, _I, _P, _P]),
        "orbit_fct_linop_csr": (_I, [_P, _P, _P, _P, _P, C.POINTER(_I)]),
        "orbit_fct_linop_destroy": (_I, [_P]),
        "orbit_bench_halfsweep": (_I, [_P, _P, _I, _I, _I, _I, C.POINTER(_D), C.POINTER(_D)]),
        "orbit_spmat_create": (_I, [_P, _I, _P, _P, _P, C.POINTER(_P)]),
        "orbit_spmat_apply_dev": (_I, [_P, _P, _I, _D, _P, _P]),
        "orbit_spmat_destroy": (_I, [_P]),
        "orbit_fct_grid_create": (_I, [_P, C.POINTER(GridMetrics), C.POINTER(_P)]),
        "orbit_fct_grid_destroy": (_I, [_P]),
        "orbit_fct_linop_create_dev": (_I, [_P, _P, _P, _D, C.POINTER(_P)]),
        "orbit_opset_check": (_I, [_P, _P, C.POINTER(C.c_longlong)]),
        "orbit_opset_autotune": (_I, [_P, _P, _I, C.POINTER(SweepShape), C.POINTER(_D)]),
        "orbit_opset_get_shapes": (_I, [_P, C.POINTER(SweepShape)]),
        "orbit_opset_set_shapes": (_I, [_P, C.POINTER(SweepShape)]),
        "orbit_transport_assemble": (_I, [_P, _P, _P]),
        "orbit_lateral_inflow": (_I, [_P, _P, _P]),
        "orbit_opset_set_relaxation": (_I, [_P, _P, _D]),
        "orbit_opset_create_grids": (_I, [_P, _I, _I, _D, _P, _P, C.POINTER(_P)]),
        "orbit_lut_create": (_I, [_P, _I, _P, _P, _I, _P, C.POINTER(_P)]),
        "orbit_lut_query_dev": (_I, [_P, _P, C.c_longlong, _P, _P, _P, _P]),
        "orbit_lut_destroy": (_I, [_P]),
        "orbit_species_diagonals": (_I, [_P, _I, _I, _I, _P, _P, _P, _P]),
        "orbit_fct_source_batch_dev": (_I, [_P, _P, _I, _P, C.c_longlong, _P, _D, _I, _P, C.c_longlong,
                                            _P, C.c_longlong]),
        "orbit_forward_fct_dev": (_I, [_P, _P, _P, _P, _I, _P, _P, _P, _I, C.POINTER(Opts), _P, _P]),
        "orbit_forward_fct_passes_dev": (_I, [_P, _P, _P, _P, _I, _P, _P, _P, _I, C.POINTER(Opts), _P, _P, _P]),
    }


def _declare(lib):
    for name, (res, args) in signatures().items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Load liborbit_hip.so from the package directory (raises if absent).

    torch is imported first: it carries its own HIP runtime, and the process
    must load that one before the library resolves its HIP dependency, or
    torch's later device initialisation finds no GPU."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OrbitLibraryMissing(
                    f"{LIB_PATH} is not built; run `python -m orbit_amd.build` "
                    "(the HIP path has no CPU fallback)")
            import torch  # noqa: F401  (load order of the HIP runtime, see above)
            h = C.CDLL(LIB_PATH)
            _declare(h)
            _lib = h
    return _lib


def last_error() -> str:
    """Message of the last error on the calling thread."""
    msg = lib().orbit_last_error()
    return msg.decode() if msg else ""


def check(rc, allow_warn=True):
    if rc == ORBIT_OK or (allow_warn and rc == ORBIT_WARN_NOT_CONVERGED):
        return rc
    msg = lib().orbit_last_error()
    raise OrbitError(rc, msg.decode() if msg else "")


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().orbit_opts_default(C.byref(o))
    for k, v in kw.items():
        if v is not None:
            setattr(o, k, v)
    return o


def dev_f64(t, device: int, name: str, last: int = None, rows: int = None):
    """t itself when it is a torch float64 tensor on cuda:device (with last
    dimension `last` and `rows` leading rows when given); else TypeError or
    ValueError before any pointer reaches a kernel (a host pointer or a
    narrower dtype would be read out of bounds on the device)."""
    import torch
    if not torch.is_tensor(t):
        raise TypeError(f"{name} must be a torch tensor on cuda:{device}")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"{name} must live on cuda:{device}, got {t.device}")
    if last is not None and (t.dim() == 0 or t.shape[-1] != last):
        raise ValueError(f"{name} must have last dimension {last}, got shape {tuple(t.shape)}")
    if rows is not None and (t.dim() < 2 or t.shape[0] != rows):
        raise ValueError(f"{name} must have {rows} rows, got shape {tuple(t.shape)}")
    return t


def ptr(a) -> int:
    """Address of a contiguous numpy array or torch tensor."""
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
