"""Loader and ctypes signatures of the HIP library (include/skymodel.h).

There is no CPU fallback: if libskymodel_hip.so is missing or cannot be loaded
this raises, so nothing silently runs elsewhere.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

from .srclist import SmComponent, ShapeletCoeff

LIB_PATH = Path(__file__).resolve().parent / "libskymodel_hip.so"
_lib = None


class SkyModelError(RuntimeError):
    """A C-ABI entry point returned an error string (the reference's GpuError::Kernel)."""


class UVW(C.Structure):
    _fields_ = [("u", C.c_double), ("v", C.c_double), ("w", C.c_double)]


class LmnRime(C.Structure):
    _fields_ = [("l", C.c_double), ("m", C.c_double), ("n", C.c_double)]


class GaussianParams(C.Structure):
    _fields_ = [("maj", C.c_double), ("min", C.c_double), ("pa", C.c_double)]


class Addresses(C.Structure):
    _fields_ = [
        ("num_freqs", C.c_int),
        ("num_vis", C.c_int),
        ("num_baselines", C.c_int),
        ("d_freqs", C.c_void_p),
        ("d_shapelet_basis_values", C.c_void_p),
        ("num_unique_beam_freqs", C.c_int),
        ("d_tile_map", C.c_void_p),
        ("d_freq_map", C.c_void_p),
        ("d_tile_index_to_unflagged_tile_index_map", C.c_void_p),
    ]


class Points(C.Structure):
    _fields_ = [
        ("num_power_laws", C.c_int), ("power_law_lmns", C.c_void_p), ("power_law_fds", C.c_void_p),
        ("power_law_sis", C.c_void_p),
        ("num_curved_power_laws", C.c_int), ("curved_power_law_lmns", C.c_void_p),
        ("curved_power_law_fds", C.c_void_p), ("curved_power_law_sis", C.c_void_p),
        ("curved_power_law_qs", C.c_void_p),
        ("num_lists", C.c_int), ("list_lmns", C.c_void_p), ("list_fds", C.c_void_p),
    ]


class Gaussians(C.Structure):
    _fields_ = [
        ("num_power_laws", C.c_int), ("power_law_lmns", C.c_void_p), ("power_law_fds", C.c_void_p),
        ("power_law_sis", C.c_void_p), ("power_law_gps", C.c_void_p),
        ("num_curved_power_laws", C.c_int), ("curved_power_law_lmns", C.c_void_p),
        ("curved_power_law_fds", C.c_void_p), ("curved_power_law_sis", C.c_void_p),
        ("curved_power_law_qs", C.c_void_p), ("curved_power_law_gps", C.c_void_p),
        ("num_lists", C.c_int), ("list_lmns", C.c_void_p), ("list_fds", C.c_void_p), ("list_gps", C.c_void_p),
    ]


class Shapelets(C.Structure):
    _fields_ = [
        ("num_power_laws", C.c_int), ("power_law_lmns", C.c_void_p), ("power_law_fds", C.c_void_p),
        ("power_law_sis", C.c_void_p), ("power_law_gps", C.c_void_p), ("power_law_shapelet_uvs", C.c_void_p),
        ("power_law_shapelet_coeffs", C.c_void_p), ("power_law_num_shapelet_coeffs", C.c_void_p),
        ("num_curved_power_laws", C.c_int), ("curved_power_law_lmns", C.c_void_p),
        ("curved_power_law_fds", C.c_void_p), ("curved_power_law_sis", C.c_void_p),
        ("curved_power_law_qs", C.c_void_p), ("curved_power_law_gps", C.c_void_p),
        ("curved_power_law_shapelet_uvs", C.c_void_p), ("curved_power_law_shapelet_coeffs", C.c_void_p),
        ("curved_power_law_num_shapelet_coeffs", C.c_void_p),
        ("num_lists", C.c_int), ("list_lmns", C.c_void_p), ("list_fds", C.c_void_p), ("list_gps", C.c_void_p),
        ("list_shapelet_uvs", C.c_void_p), ("list_shapelet_coeffs", C.c_void_p),
        ("list_num_shapelet_coeffs", C.c_void_p),
    ]


class RADec(C.Structure):
    _fields_ = [("ra", C.c_double), ("dec", C.c_double)]


class IonoConsts(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta", C.c_double), ("gain", C.c_double)]


class SmObsParams(C.Structure):
    _fields_ = [
        ("num_unflagged_tiles", C.c_int32),
        ("unflagged_tile_xyzs", C.c_void_p),
        ("num_flagged_tiles", C.c_int32),
        ("flagged_tiles", C.c_void_p),
        ("num_freqs", C.c_int32),
        ("freqs", C.c_void_p),
        ("phase_centre_ra", C.c_double),
        ("phase_centre_dec", C.c_double),
        ("array_longitude", C.c_double),
        ("array_latitude", C.c_double),
        ("beam_type", C.c_int32),
        ("dipole_delays", C.c_void_p),
        ("dipole_gains", C.c_void_p),
        ("pols", C.c_int32),
        ("dut1", C.c_double),
        ("apply_precession", C.c_int32),
    ]


class HADec(C.Structure):
    _fields_ = [("ha", C.c_double), ("dec", C.c_double)]


class XYZ(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class SmPrecessionInfo(C.Structure):
    _fields_ = [
        ("rotation_matrix", (C.c_double * 3) * 3),
        ("hadec_j2000", HADec),
        ("lmst", C.c_double),
        ("lmst_j2000", C.c_double),
        ("array_latitude_j2000", C.c_double),
    ]


class SmUvfitsParams(C.Structure):
    _fields_ = [
        ("num_timesteps", C.c_int32),
        ("num_baselines", C.c_int32),
        ("num_chans", C.c_int32),
        ("num_tiles", C.c_int32),
        ("tile_names", C.c_void_p),
        ("tile_xyzs", C.c_void_p),
        ("baseline_tiles", C.c_void_p),
        ("first_freq_hz", C.c_double),
        ("freq_res_hz", C.c_double),
        ("phase_centre_ra", C.c_double),
        ("phase_centre_dec", C.c_double),
        ("array_longitude", C.c_double),
        ("array_latitude", C.c_double),
        ("array_height", C.c_double),
        ("dut1", C.c_double),
        ("apply_precession", C.c_int32),
        ("int_time_s", C.c_double),
        ("first_gps_seconds", C.c_double),
        ("obs_name", C.c_char_p),
    ]


class SmInfo(C.Structure):
    _fields_ = [
        ("num_baselines", C.c_int32),
        ("num_freqs", C.c_int32),
        ("num_unflagged_tiles", C.c_int32),
        ("num_unique_tiles", C.c_int32),
        ("num_points", C.c_int32),
        ("num_gaussians", C.c_int32),
        ("num_shapelets", C.c_int32),
        ("num_kernel_launches", C.c_int32),
        ("kernel_ms", C.c_double),
        ("kernel_ms_points", C.c_double),
        ("kernel_ms_gaussians", C.c_double),
        ("kernel_ms_shapelets", C.c_double),
        ("kernel_ms_factored", C.c_double),
        ("num_factored_launches", C.c_int32),
    ]


# Every symbol the library exports through include/*.h, with its ctypes signature.
_P = C.c_void_p
_D = C.c_double
_I = C.c_int32
SIGNATURES = {
    "model_points": [_P, _P, _P, _P, _P],
    "model_gaussians": [_P, _P, _P, _P, _P],
    "model_shapelets": [_P, _P, _P, _P, _P],
    "model_autos": [C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P, _P],
    "sm_model_points_f64": [_P, _P, _P, _P, _P, _P],
    "sm_model_gaussians_f64": [_P, _P, _P, _P, _P, _P],
    "sm_model_shapelets_f64": [_P, _P, _P, _P, _P, _P],
    "sm_model_points_noalloc": [_P, _P, _P, _P, _P, _P],
    "sm_new": [C.POINTER(SmObsParams), _P, _P, _I, C.POINTER(_P)],
    "sm_update_with_a_source": [_P, _P, _I, _D, _D],
    "sm_model_timestep_with": [_P, _D, _D, _P, _P, _P],
    "sm_model_timestep_with_device": [_P, _D, _D, _P, _P, _P],
    "sm_model_timestep_autos_with": [_P, _D, _D, _P, _P],
    "sm_get_info": [_P, C.POINTER(SmInfo)],
    "sm_set_timing": [_P, _I],
    "sm_synchronize": [],
    "sm_set_direct_variant": [_I],
    "sm_set_class_table_limit": [C.c_int64],
    "xyzs_to_uvws": [_P, _P, _P, RADec, C.c_int, C.c_int, C.c_int],
    "rotate": [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P],
    "add_model": [_P, _P, IonoConsts, _P, _P, C.c_int, C.c_int, C.c_int],
    "sm_get_flux_params": [_P, _I, _I, _I, _I, _P, C.POINTER(_D), C.POINTER(_D)],
    "sm_analytic_beam": [_I, _P, _P, _I, _P, _I, _P, _P, _P],
    "sm_model_timestep_factored": [_P, _D, _D, _I, _P, _P],
    "sm_selftest_mfma_layout": [_P, _P, _P],
    "sm_set_factored_segment": [C.c_int32],
    "sm_set_factored_kernel": [C.c_int32],
    "sm_surrogate_new": [C.c_uint64, C.POINTER(_P)],
    "sm_surrogate_new_shaped": [C.c_uint64, _I, _I, C.POINTER(_P)],
    "sm_surrogate_shape": [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(C.c_int64)],
    "sm_surrogate_get_params": [_P, _P],
    "sm_surrogate_set_params": [_P, _P],
    "sm_surrogate_samples": [C.c_uint64, C.c_int64, _I, _I, _P, _P, _P],
    "sm_surrogate_gradients": [_P, _I, _P, _P, C.POINTER(_D), _P],
    "sm_surrogate_train_step": [_P, _I, _P, _P, C.c_float, C.POINTER(_D)],
    "sm_surrogate_forward": [_P, _I, _P, _I, _P],
    "sm_surrogate_beam": [_P, _I, _P, _P, _I, _P, _I, _P, _P, _D, _P, C.POINTER(_I)],
    "sm_set_beam_surrogate": [_P, _P, _D, C.POINTER(_I)],
    "sm_precess_time": [_D, _D, _D, _D, _D, _D, C.POINTER(SmPrecessionInfo)],
    "sm_precess_xyz": [C.POINTER(SmPrecessionInfo), _P, _P, _I],
    "sm_lst_uvws_latitude": [_P, _D, C.POINTER(_D), C.POINTER(_D), _P],
    "sm_model_timestep_epoch_with": [_P, _D, _P, _P, _P],
    "sm_model_timestep_epoch_with_device": [_P, _D, _P, _P],
    "sm_model_timestep_autos_epoch_with": [_P, _D, _P, _P],
    "sm_uvfits_create": [C.c_char_p, C.POINTER(SmUvfitsParams), C.POINTER(_P)],
    "sm_uvfits_write_timestep": [_P, _D, _P, _P],
    "sm_uvfits_finish": [_P],
}
# Entry points that return a value rather than an error string.
VALUE_RETURN = {"sm_get_lmst": ([_D, _D, _D], C.c_double), "sm_surrogate_num_params": ([], C.c_int32)}
VOID_RETURN = {"sm_free", "sm_surrogate_free"}


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SkyModelError(
                f"{LIB_PATH} is missing: build it with `python -m mwa_hyperdrive_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = C.c_char_p
        for name in VOID_RETURN:
            getattr(L, name).argtypes = [_P]
            getattr(L, name).restype = None
        for name, (args, res) in VALUE_RETURN.items():
            getattr(L, name).argtypes = args
            getattr(L, name).restype = res
        _lib = L
    return _lib


def check(err) -> None:
    if err is not None:
        raise SkyModelError(err.decode() if isinstance(err, bytes) else str(err))


def call(name, *args) -> None:
    check(getattr(lib(), name)(*args))
