"""ctypes binding of libswalbe_hip.so (the C ABI of include/swalbe_hip.h).

The library is required: importing the package without it raises, and every
op runs on the GPU through it.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("SWALBE_HIP_LIB", os.path.join(PKG_ROOT, "lib", "libswalbe_hip.so"))

OK, EDOMAIN, EINVAL, ENOMEM, EHIP = 0, 1, 2, 3, 4
F64, F32 = 0, 1
SIMPLE, THERMAL = 0, 1

FIELDS = {
    "fout": 0, "ftemp": 1, "feq": 2, "height": 3, "velx": 4, "vely": 5, "vsq": 6,
    "pressure": 7, "Fx": 8, "Fy": 9, "slipx": 10, "slipy": 11, "hgradpx": 12, "hgradpy": 13,
    "kbtx": 14, "kbty": 15, "theta": 16,
}
DIST_FIELDS = ("fout", "ftemp", "feq")

FIELDS_1D = {
    "fout": 0, "ftemp": 1, "feq": 2, "height": 3, "vel": 4, "pressure": 5, "F": 6, "slip": 7,
    "hgradp": 8, "kbt": 9, "theta": 10, "gamma": 11, "gradgamma": 12, "border1": 13,
    "border2": 14, "interior": 15,
}


class SwalbeError(RuntimeError):
    """A HIP runtime, allocation or argument error reported by the library;
    `status` is the library's status code (ENOMEM, EINVAL, EHIP)."""

    def __init__(self, msg: str, status: int = 0):
        super().__init__(msg)
        self.status = status


class DomainError(ValueError):
    """Raised where the reference raises DomainError (src/pressure.jl:101-107)."""


class Params(C.Structure):
    """swalbe_params = Taumucs{T} (src/initialize.jl:43-60)."""
    _fields_ = [
        ("Tmax", C.c_int64), ("tdump", C.c_int64), ("tau", C.c_double), ("cs", C.c_double),
        ("mu", C.c_double), ("delta", C.c_double), ("kbt", C.c_double), ("gamma", C.c_double),
        ("n", C.c_int32), ("m", C.c_int32), ("hmin", C.c_double), ("hcrit", C.c_double),
        ("theta", C.c_double), ("g", C.c_double),
    ]


class StepOpts(C.Structure):
    _fields_ = [
        ("use_theta_field", C.c_int32), ("thermal", C.c_int32), ("body", C.c_int32),
        ("materialize", C.c_int32), ("seed", C.c_uint64), ("t0", C.c_int64),
        ("body_ax", C.c_double), ("body_ay", C.c_double), ("body_fac", C.c_double),
        ("wet_thresh", C.c_double), ("dh_series", C.POINTER(C.c_double)),
        ("wet_series", C.POINTER(C.c_int64)), ("mass_series", C.POINTER(C.c_double)),
        ("tile_y", C.c_int32), ("variant", C.c_int32), ("nontemporal", C.c_int32),
        ("no_fp64_guard", C.c_int32), ("snap_every", C.c_int64), ("snap_series", C.POINTER(C.c_double)),
        ("maxh_series", C.POINTER(C.c_double)), ("area_series", C.POINTER(C.c_double)),
        ("energy_series", C.POINTER(C.c_double)), ("fluid_series", C.POINTER(C.c_uint8)),
        ("surf_thresh", C.c_double), ("surf_theta_field", C.c_int32), ("final_outputs_only", C.c_int32),
    ]


class StepOpts1D(C.Structure):
    """swalbe1d_step_opts: options of the fused 1D time loop."""
    _fields_ = [
        ("use_theta_field", C.c_int32), ("use_gamma_field", C.c_int32), ("grad_gamma", C.c_int32),
        ("thermal", C.c_int32), ("body", C.c_int32), ("bound", C.c_int32),
        ("materialize", C.c_int32), ("variant", C.c_int32), ("seed", C.c_uint64),
        ("t0", C.c_int64), ("alpha", C.c_double), ("body_fac", C.c_double),
        ("wet_thresh", C.c_double), ("dh_series", C.POINTER(C.c_double)),
        ("wet_series", C.POINTER(C.c_int64)), ("mass_series", C.POINTER(C.c_double)),
        ("snap_every", C.c_int64), ("snap_series", C.POINTER(C.c_double)),
        ("threads", C.c_int32), ("pad_", C.c_int32),
    ]


_lib = None
_V = C.c_void_p
_D = C.c_double
_I = C.c_int32


class FilmInit(C.Structure):
    """swalbe_film_init: initial height spec of a film (flat, noisy, one or two caps)."""
    _fields_ = [
        ("kind", C.c_int32), ("thermal", C.c_int32), ("body", C.c_int32), ("clamp", C.c_int32),
        ("seed", C.c_uint64), ("base", C.c_double), ("amp", C.c_double),
        ("precursor", C.c_double), ("href", C.c_double),
        ("r1", C.c_double), ("theta1", C.c_double), ("c1x", C.c_double), ("c1y", C.c_double),
        ("r2", C.c_double), ("theta2", C.c_double), ("c2x", C.c_double), ("c2y", C.c_double),
        ("body_ax", C.c_double), ("body_ay", C.c_double), ("body_fac", C.c_double),
    ]


def _sig(L, name, args):
    fn = getattr(L, name)
    fn.restype = C.c_int
    fn.argtypes = args


def lib():
    """Loads libswalbe_hip.so once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libswalbe_hip.so not found at {LIB_PATH}; build it with "
            "`python swalbe.jl_amd/build.py` (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    L.swalbe_last_error.restype = C.c_char_p
    L.swalbe_last_error.argtypes = []
    L.swalbe_version.restype = C.c_int
    _sig(L, "swalbe_device_count", [C.POINTER(C.c_int32)])
    _sig(L, "swalbe_create", [_I, _I, _I, _I, C.POINTER(_V)])
    _sig(L, "swalbe_destroy", [_V])
    _sig(L, "swalbe_set_reference_height", [_V, _D, _D])
    _sig(L, "swalbe_upload", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe_download", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe_fill", [_V, _I, _D])
    _sig(L, "swalbe_copy", [_V, _I, _I])
    _sig(L, "swalbe_download_raw", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe_upload_raw", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe_get_reference_height", [_V, C.POINTER(_D), C.POINTER(_D)])
    _sig(L, "swalbe_synchronize", [_V])
    _sig(L, "swalbe_set_stream", [_V, _V])
    _sig(L, "swalbe_equilibrium", [_V, _D])
    _sig(L, "swalbe_bgk_and_stream", [_V, _D])
    _sig(L, "swalbe_moments", [_V])
    _sig(L, "swalbe_filmpressure", [_V, _D, _D, _I, _I, _D, _D, _I, _I])
    _sig(L, "swalbe_hgradp", [_V])
    _sig(L, "swalbe_slippage", [_V, _D, _D])
    _sig(L, "swalbe_thermal", [_V, _D, _D, _D, C.c_uint64, C.c_uint64])
    _sig(L, "swalbe_inclination", [_V, _D, _D, _D])
    _sig(L, "swalbe_force_sum", [_V, _I])
    _sig(L, "swalbe_time_loop", [_V, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts)])
    _sig(L, "swalbe_time_loop_timed",
         [_V, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts), C.POINTER(_D)])
    _sig(L, "swalbe_reduce", [_V, _I, _D, C.POINTER(_D)])
    _sig(L, "swalbe_init_random", [_V, _I, C.c_uint64, _D, _D])
    _sig(L, "swalbe_init_height", [_V, C.POINTER(FilmInit)])
    _sig(L, "swalbe_measure", [_V, _D, _I, _D, _I, C.POINTER(_D), _V, C.c_size_t])
    _sig(L, "swalbe_measure_ptr", [_I, _V, _V, _D, _I, _I, _D, _V, _V, _V])
    _sig(L, "swalbe_measure_ptr_noalloc", [_I, _V, _V, _D, _I, _I, _D, _V, _V, _V])
    _sig(L, "swalbe_grad_simple_ptr", [_I, _V, _V, _V, _I, _I, _V])
    _sig(L, "swalbe_device_memory", [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)])
    _sig(L, "swalbe_batch_create", [_I, C.c_size_t, _I, C.POINTER(_V)])
    _sig(L, "swalbe_batch_destroy", [_V])
    _sig(L, "swalbe_batch_film_bytes", [_I, _I, _D, C.POINTER(C.c_size_t)])
    _sig(L, "swalbe_batch_add", [_V, C.POINTER(Params), _I, _I, C.POINTER(FilmInit), C.c_int64,
                                 C.POINTER(_I)])
    _sig(L, "swalbe_batch_remove", [_V, _I])
    _sig(L, "swalbe_batch_step", [_V, C.c_int64])
    _sig(L, "swalbe_batch_step_timed", [_V, C.c_int64, C.POINTER(_D)])
    _sig(L, "swalbe_batch_set_graph", [_V, _I])
    _sig(L, "swalbe_batch_set_nontemporal", [_V, _I])
    _sig(L, "swalbe_batch_set_final_outputs_only", [_V, _I])
    _sig(L, "swalbe_batch_set_autotune", [_V, _I])
    _sig(L, "swalbe_batch_set_tile_y", [_V, _I])
    _sig(L, "swalbe_batch_tune_shape", [C.POINTER(Params), _I, _I, C.POINTER(FilmInit),
                                        C.POINTER(_I), C.POINTER(_D), C.POINTER(_D)])
    _sig(L, "swalbe_batch_film_tile_y", [_V, _I, C.POINTER(_I)])
    _sig(L, "swalbe_batch_sync", [_V])
    _sig(L, "swalbe_batch_wait", [_V, C.c_int64])
    _sig(L, "swalbe_batch_film_status", [_V, _I, C.POINTER(C.c_int64), C.POINTER(_I)])
    _sig(L, "swalbe_batch_download", [_V, _I, _I, _V, C.c_size_t])
    _sig(L, "swalbe_serve_create", [C.c_size_t, _I, _I, _I, C.POINTER(_V)])
    _sig(L, "swalbe_serve_destroy", [_V])
    _sig(L, "swalbe_serve_submit", [_V, C.POINTER(Params), _I, _I, C.POINTER(FilmInit), C.c_int64,
                                    C.POINTER(C.c_int64)])
    _sig(L, "swalbe_serve_poll", [_V, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)])
    _sig(L, "swalbe_serve_peek", [_V, C.POINTER(C.c_int64), C.POINTER(_I), C.POINTER(_I)])
    _sig(L, "swalbe_serve_take", [_V, C.POINTER(_D), C.c_size_t])
    _sig(L, "swalbe_batch_stats", [_V, C.POINTER(_I), C.POINTER(C.c_int64), C.POINTER(C.c_size_t),
                                   C.POINTER(C.c_size_t)])
    _sig(L, "swalbe_equilibrium_ptr", [_I, _V, _V, _V, _V, _V, _D, _I, _I, _V])
    _sig(L, "swalbe_bgk_and_stream_ptr", [_I, _V, _V, _V, _V, _V, _D, _I, _I, _V])
    _sig(L, "swalbe_moments_ptr", [_I, _V, _V, _V, _V, _I, _I, _V])
    _sig(L, "swalbe_filmpressure_ptr", [_I, _V, _V, _D, _D, _V, _I, _I, _D, _D, _I, _I, _I, _V])
    _sig(L, "swalbe_grad_ptr", [_I, _V, _V, _V, _V, _I, _I, _V])
    _sig(L, "swalbe_laplacian_ptr", [_I, _V, _V, _D, _I, _I, _V])
    _sig(L, "swalbe_hgradp_ptr", [_I, _V, _V, _V, _V, _I, _I, _V])
    _sig(L, "swalbe_slippage_ptr", [_I, _V, _V, _V, _V, _V, _D, _D, _I, _I, _V])
    _sig(L, "swalbe_thermal_ptr", [_I, _V, _V, _V, _D, _D, _D, C.c_uint64, C.c_uint64, _I, _I, _V])
    _sig(L, "swalbe_inclination_ptr", [_I, _V, _V, _V, _D, _D, _D, _I, _I, _V])
    _sig(L, "swalbe_malloc", [C.POINTER(_V), C.c_size_t])
    _sig(L, "swalbe_free", [_V])
    _sig(L, "swalbe_memcpy_h2d", [_V, _V, C.c_size_t])
    _sig(L, "swalbe_memcpy_d2h", [_V, _V, C.c_size_t])
    _sig(L, "swalbe_host_normals", [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64, C.POINTER(_D), C.POINTER(_D)])
    _sig(L, "swalbe_host_uniforms", [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int64, C.POINTER(_D)])
    _sig(L, "swalbe1d_create", [_I, _I, _I, C.POINTER(_V)])
    _sig(L, "swalbe1d_destroy", [_V])
    _sig(L, "swalbe1d_length", [_V, C.POINTER(_I)])
    _sig(L, "swalbe1d_upload", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe1d_download", [_V, _I, _V, C.c_size_t])
    _sig(L, "swalbe1d_fill", [_V, _I, _D])
    _sig(L, "swalbe1d_synchronize", [_V])
    _sig(L, "swalbe1d_equilibrium", [_V, _D])
    _sig(L, "swalbe1d_bgk_and_stream", [_V, _D, _I])
    _sig(L, "swalbe1d_moments", [_V])
    _sig(L, "swalbe1d_filmpressure", [_V, _D, _D, _I, _I, _D, _D, _I, _I, _I])
    _sig(L, "swalbe1d_grad_gamma", [_V, _I, _D])
    _sig(L, "swalbe1d_hgradp", [_V])
    _sig(L, "swalbe1d_slippage", [_V, _D, _D])
    _sig(L, "swalbe1d_thermal", [_V, _D, _D, _D, C.c_uint64, C.c_uint64])
    _sig(L, "swalbe1d_inclination", [_V, _D, _D])
    _sig(L, "swalbe1d_force_sum", [_V, _I, _I])
    _sig(L, "swalbe1d_reduce", [_V, _I, _D, _I, C.POINTER(_D)])
    _sig(L, "swalbe1d_time_loop", [_V, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts1D)])
    _sig(L, "swalbe1d_time_loop_timed",
         [_V, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts1D), C.POINTER(_D)])
    _sig(L, "swalbe1d_ensemble_time_loop",
         [C.POINTER(_V), _I, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts1D)])
    _sig(L, "swalbe1d_ensemble_time_loop_timed",
         [C.POINTER(_V), _I, C.POINTER(Params), C.c_int64, C.POINTER(StepOpts1D), C.POINTER(_D)])
    _sig(L, "swalbe1d_resident_max_length", [_I, _I, C.POINTER(_I)])
    _sig(L, "swalbe1d_grad_ptr", [_I, _V, _V, _V, _I, _V])
    _sig(L, "swalbe1d_laplacian_ptr", [_I, _V, _V, _I, _V])
    L.swalbe_host_cospi.restype = C.c_double
    L.swalbe_host_cospi.argtypes = [C.c_double]
    _lib = L
    return L


def check(rc: int, what: str = "") -> None:
    if rc == OK:
        return
    msg = lib().swalbe_last_error().decode(errors="replace")
    if rc == EDOMAIN:
        raise DomainError(msg)
    raise SwalbeError(f"{what}: {msg} (status {rc})", rc)


def device_count() -> int:
    n = C.c_int32(0)
    check(lib().swalbe_device_count(C.byref(n)), "swalbe_device_count")
    return n.value
