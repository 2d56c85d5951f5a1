"""ctypes binding of libletkf_engine.so (the C ABI of include/letkf_engine.h).

The engine has no CPU fallback: importing this module without the built
library, or calling it without a HIP device, raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libletkf_engine.so")

LETKF_OK = 0
STATUS_NAMES = {0: "LETKF_OK", -1: "LETKF_ERR_ARG", -2: "LETKF_ERR_HIP",
                -3: "LETKF_ERR_EIG", -4: "LETKF_ERR_CAPACITY",
                -5: "LETKF_ERR_NONFINITE"}

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_FP32_PANEL, PREC_FP32X3, PREC_FP32_DUAL64 = 0, 1, 2, 3, 4, 5
PREC_BF16X3_TRI32 = 6
FORM_PLAIN, FORM_BATCHED, FORM_BATCHED_QL = 0, 1, 2

# every symbol include/letkf_engine.h declares
EXPORTS = ("letkf_params_default", "letkf_create", "letkf_destroy",
           "letkf_da_main", "letkf_da_main_device", "letkf_core_batched",
           "letkf_bin_obs", "letkf_aoei", "letkf_da_adjoint_device",
           "letkf_set_params", "letkf_get_stats", "letkf_last_error", "letkf_trans_xtoy_device",
           "letkf_setobs_default", "letkf_set_obs_device", "letkf_da_main_plan", "letkf_da_main_enqueue",
           "letkf_da_main_status", "letkf_autotune", "letkf_set_tuning", "letkf_get_tuning",
           "letkf_ensemble_mesp_device", "letkf_grid_pack_device", "letkf_lpfgm_core_batched",
           "letkf_da_main_lpfgm_device", "letkf_da_main_lpfgm", "letkf_get_nobsl",
           "letkf_core_batched_noalloc", "letkf_rational_table_entry")


class LpfgmCfg(C.Structure):
    """letkf_lpfgm_cfg_t: iswitch_da, nmonte, RM, DP, cov_infl_gm, draws."""
    _fields_ = [(n, C.c_int32) for n in ("mode", "nmonte", "rm_sr", "dp")] + [
        ("cov_infl_gm", C.c_double), ("rand", C.POINTER(C.c_double))]


class Params(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nmem", "nlon", "nlat", "nlev", "nij1", "nv3d", "nv2d", "iv3d_t",
        "iv3d_s", "id_z_obs", "real_kind", "roff", "precision", "form")] + [
        (n, C.c_double) for n in (
            "pi", "re", "cov_infl_mul", "alpha_rtpp", "alpha_rtps",
            "sigma_obs", "sigma_obsv", "dist_zero", "dist_zerov", "tmin",
            "tmax", "smin", "smax")] + [("stage_timing", C.c_int32), ("workspace_bytes", C.c_int64)]


class Grid(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in (
        "lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1", "depth1")]


class Obs(C.Structure):
    _fields_ = [("nobs", C.c_int32), ("nobsgrd", C.c_void_p),
                ("elm", C.c_void_p)] + [(n, C.c_void_p) for n in (
                    "lon", "lat", "lev", "err", "dep", "hdxf")]


class XtoyGrid(C.Structure):
    """letkf_xtoy_grid_t: the global grid and element constants of H(x)."""
    _fields_ = [(n, C.c_int32) for n in (
        "nlon", "nlat", "nlev", "nv3d", "nv2d", "real_kind", "id_u_obs", "id_v_obs", "id_t_obs",
        "id_s_obs", "id_z_obs", "iv3d_u", "iv3d_v", "iv3d_t", "iv3d_s", "iv2d_z")] + [
        ("hnosea", C.c_double), ("undef", C.c_double)] + [
        (n, C.c_void_p) for n in ("lon", "lat", "depth", "fsm")]


class SetObsCfg(C.Structure):
    """letkf_setobs_cfg_t: settings of set_obs."""
    _fields_ = [(n, C.c_int32) for n in (
        "id_u_obs", "id_v_obs", "id_t_obs", "id_s_obs", "id_z_obs", "lge_io", "aoei", "nslots",
        "nbslot")] + [(n, C.c_double) for n in ("sigma_obst", "undef", "gross_error")] + [
        ("qc_min", C.c_double * 9), ("qc_max", C.c_double * 9)]


class ObsSet(C.Structure):
    """letkf_obs_set_t: device outputs of letkf_set_obs_device."""
    _fields_ = [(n, C.c_void_p) for n in (
        "nobsgrd", "elm", "ins", "idx", "idy", "lon", "lat", "lev", "dat", "err", "dep", "hxfmean",
        "hxfsprd", "hdxf")] + [(n, C.c_int32) for n in ("nobs_out", "ndel_proj", "ndel_gross")]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "points", "active_points", "local_obs", "dual_points", "max_nobsl")] + [
        (n, C.c_double) for n in (
            "ms_total", "ms_search", "ms_gram", "ms_eig", "ms_transform",
            "gram_flops", "transform_bytes")] + [("chunks", C.c_int64), ("eig_flops", C.c_double),
                                                 ("ms_eigk", C.c_double * 25), ("n_eigk", C.c_int64 * 25),
                                                 ("rotations", C.c_int64), ("fallback_slots", C.c_int64),
                                                 ("bucket_slots", C.c_int64 * 5)]


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load the engine library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C sbpom_letkf_v2_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.letkf_params_default.argtypes = [C.POINTER(Params), C.c_int32]
        L.letkf_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.letkf_destroy.argtypes = [vp]
        L.letkf_da_main.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp, vp, vp, vp]
        L.letkf_da_main_device.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp, vp, vp, vp, vp]
        if hasattr(L, "letkf_core_batched"):
            L.letkf_core_batched.argtypes = [vp, C.c_int32, vp, vp, vp, vp, vp, vp, vp]
        if hasattr(L, "letkf_bin_obs"):
            L.letkf_bin_obs.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, C.c_int32, vp,
                                        vp, vp, vp, vp, vp, vp]
        if hasattr(L, "letkf_aoei"):
            L.letkf_aoei.argtypes = [vp, C.c_int32, vp, vp, vp]
        L.letkf_da_main_plan.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp]
        L.letkf_da_main_enqueue.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp, vp, vp, vp, vp]
        L.letkf_da_main_status.argtypes = [vp, vp]
        L.letkf_autotune.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp, vp, vp, vp,
                                     C.POINTER(C.c_double), vp]
        L.letkf_set_tuning.argtypes = [vp, C.POINTER(C.c_int32)]
        L.letkf_get_tuning.argtypes = [vp, C.POINTER(C.c_int32)]
        L.letkf_da_adjoint_device.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), vp, vp, vp, vp,
                                              vp, vp, vp, vp]
        L.letkf_set_params.argtypes = [vp, C.POINTER(Params)]
        L.letkf_trans_xtoy_device.argtypes = [C.POINTER(XtoyGrid), C.c_int32, vp, vp, vp, vp, vp, vp,
                                              C.c_int32, vp, vp, C.POINTER(C.c_int64), vp, vp, vp]
        L.letkf_setobs_default.argtypes = [C.POINTER(SetObsCfg)]
        L.letkf_set_obs_device.argtypes = ([vp, C.POINTER(SetObsCfg), C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(C.c_int32)] + [vp] * 11 + [C.POINTER(ObsSet)])
        L.letkf_get_stats.argtypes = [vp, C.POINTER(Stats)]
        L.letkf_get_nobsl.argtypes = [vp, vp, vp]
        L.letkf_rational_table_entry.argtypes = [C.c_int32, C.c_int32, vp, vp, vp]
        L.letkf_core_batched_noalloc.argtypes = [C.POINTER(Params), C.c_int32] + [vp] * 9
        i32 = C.c_int32
        L.letkf_ensemble_mesp_device.argtypes = [i32] * 6 + [C.c_double] + [vp] * 7
        L.letkf_da_main_lpfgm_device.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), C.POINTER(LpfgmCfg),
                                                 vp, vp, vp, vp, vp]
        L.letkf_da_main_lpfgm.argtypes = [vp, C.POINTER(Grid), C.POINTER(Obs), C.POINTER(LpfgmCfg), vp, vp, vp, vp]
        L.letkf_lpfgm_core_batched.argtypes = ([vp, i32, C.c_double, i32, i32, i32, C.POINTER(C.c_double), i32]
                                               + [vp] * 7)
        L.letkf_grid_pack_device.argtypes = ([i32] * 9 + [vp, C.POINTER(C.c_int64), vp, C.POINTER(C.c_int64)]
                                             + [vp] * 3)
        L.letkf_last_error.restype = C.c_char_p
        for n in EXPORTS:
            if n not in ("letkf_last_error", "letkf_setobs_default") and hasattr(L, n):
                getattr(L, n).restype = C.c_int
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != LETKF_OK:
        raise EngineError(rc, lib().letkf_last_error().decode(errors="replace"))
