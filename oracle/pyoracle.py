"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference
build (_ref/libletkf_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the sbpom_letkf_v2_amd package.

The oracle restates letkf/run/src/common_da.f90:22-533,
common_letkf.f90:37-210, common.f90:79-168,992-1050 and
common_obs.f90:266-400 in C (letkf_oracle_body.inc cites each routine).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libletkf_ref.so")
REF_SRC = "/root/reference/letkf/run/src"


def build(ref: bool = True) -> None:
    """Compile liboracle.so, and _ref/ where the reference sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir(REF_SRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


class OracleParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        "nmem", "nlon", "nlat", "nlev", "nij1", "nv3d", "nv2d", "iv3d_t",
        "iv3d_s", "id_z_obs", "roff", "nthreads")] + [(n, C.c_double) for n in (
        "pi", "re", "cov_infl_mul", "alpha_rtpp", "alpha_rtps", "sigma_obs",
        "sigma_obsv", "dist_zero", "dist_zerov", "tmin", "tmax", "smin",
        "smax")] + [(n, C.c_int32) for n in ("iswitch_da", "nmonte", "rm_sr", "dp")] + [
        ("cov_infl_gm", C.c_double), ("rand", C.POINTER(C.c_double))]


def _tables_struct(ct):
    P = C.POINTER(ct)
    IP = C.POINTER(C.c_int32)
    fields = [(n, P) for n in ("lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1",
                               "phi1", "depth1")]
    fields += [("nobs", C.c_int32), ("nobsgrd", IP), ("elm", IP)]
    fields += [(n, P) for n in ("olon", "olat", "olev", "oerr", "odep", "ohdxf")]
    return type("Tables", (C.Structure,), {"_fields_": fields})


class OracleXtoY(C.Structure):
    """oracle_xtoy_t: grid and element constants of Trans_XtoY."""
    _fields_ = [(n, C.c_int32) for n in (
        "nlon", "nlat", "nlev", "nv3d", "nv2d", "id_u_obs", "id_v_obs", "id_t_obs", "id_s_obs",
        "id_z_obs", "iv3d_u", "iv3d_v", "iv3d_t", "iv3d_s", "iv2d_z")] + [
        ("hnosea", C.c_double), ("undef", C.c_double)]


Tables64 = _tables_struct(C.c_double)
Tables32 = _tables_struct(C.c_float)

_lib = None
_ref = None


class OracleSetObs(C.Structure):
    """oracle_setobs_t: settings of set_obs (common_setting.f90:56-63,
    96, 101-118; common_obs.f90:35)."""
    _fields_ = [(n, C.c_int32) for n in (
        "nmem", "nlon", "nlat", "nslots", "nbslot", "id_u_obs", "id_v_obs", "id_t_obs",
        "id_s_obs", "id_z_obs", "lge_io", "aoei")] + [
        (n, C.c_double) for n in ("sigma_obst", "undef", "gross_error")] + [
        ("qc_min", C.c_double * 9), ("qc_max", C.c_double * 9)]


def setobs_config(nmem, nlon, nlat, **kw) -> OracleSetObs:
    """The reference's settings (LGE_IO off, AOEI on, one slot), overridden
    by keyword (qc_min / qc_max as 9-sequences)."""
    c = OracleSetObs(nmem=nmem, nlon=nlon, nlat=nlat, nslots=1, nbslot=1, id_u_obs=2819,
                     id_v_obs=2820, id_t_obs=3073, id_s_obs=3332, id_z_obs=2567, lge_io=0, aoei=1,
                     sigma_obst=1.0, undef=999.0, gross_error=10.0)
    lo = (-1, -5, -5, -2, -2, -2, -2, -2, -2)
    for k in range(9):
        c.qc_min[k], c.qc_max[k] = lo[k], -lo[k]
    for k, v in kw.items():
        if k in ("qc_min", "qc_max"):
            for i, x in enumerate(v):
                getattr(c, k)[i] = x
        else:
            setattr(c, k, v)
    return c


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        _lib = C.CDLL(LIB)
        for suf, ct, T in (("_f64", C.c_double, Tables64), ("_f32", C.c_float, Tables32)):
            P = C.POINTER(ct)
            DP = C.POINTER(C.c_double)
            IP = C.POINTER(C.c_int32)
            f = getattr(_lib, "oracle_distance" + suf)
            f.restype = ct
            f.argtypes = [ct] * 6
            f = getattr(_lib, "oracle_eigen" + suf)
            f.restype = C.c_int
            f.argtypes = [C.c_int, P, DP, DP]
            f = getattr(_lib, "oracle_letkf_core" + suf)
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, C.c_double, P, P, P, P, P, P]
            f = getattr(_lib, "oracle_count_nobsl" + suf)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(OracleParams), C.POINTER(T), C.c_int, C.c_int]
            f = getattr(_lib, "oracle_obs_local" + suf)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(OracleParams), C.POINTER(T), C.c_int, C.c_int,
                          C.c_int, IP, P, P, P, P]
            f = getattr(_lib, "oracle_analyse_points" + suf)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(OracleParams), C.POINTER(T), C.c_int64, IP, IP,
                          P, P, P, P, IP]
            f = getattr(_lib, "oracle_da_main" + suf)
            f.restype = C.c_int
            f.argtypes = [C.POINTER(OracleParams), C.POINTER(T), P, P, P, P, IP]
            f = getattr(_lib, "oracle_aoei" + suf)
            f.restype = None
            f.argtypes = [C.c_int, P, P, P]
            f = getattr(_lib, "oracle_bin_obs" + suf)
            f.restype = C.c_int
            f.argtypes = [C.c_int, C.c_int, P, P, C.c_int, P, P, IP, IP, IP, IP]
            f = getattr(_lib, "oracle_akima" + suf)
            f.restype = ct
            f.argtypes = [C.c_int, P, P, ct, ct]
            f = getattr(_lib, "oracle_trans_xtoy" + suf)
            f.restype = None
            f.argtypes = [C.POINTER(OracleXtoY), C.c_int, IP, IP, IP, P, P, P, P, P, P, P, P, P, P]
            f = getattr(_lib, "oracle_set_obs" + suf)
            f.restype = C.c_int
            f.argtypes = ([C.POINTER(OracleSetObs), C.c_int, IP, IP, IP] + [P] * 5 + [IP, IP, P, IP]
                          + [IP] * 4 + [P] * 9 + [IP, IP])
            f = getattr(_lib, "oracle_ensemble_mesp" + suf)
            f.restype = None
            f.argtypes = [C.c_int] * 5 + [ct] + [P] * 6
            f = getattr(_lib, "oracle_lpfgm_core" + suf)
            f.restype = C.c_int
            f.argtypes = [C.c_int] * 3 + [C.c_double] + [C.c_int] * 3 + [C.POINTER(C.c_double)] + [P] * 6
            f = getattr(_lib, "oracle_grid_pack" + suf)
            f.restype = None
            FP = C.POINTER(C.c_float)
            f.argtypes = [C.c_int] * 8 + [P] + [C.c_int64] * 3 + [P] + [C.c_int64] * 2 + [FP, FP]
    return _lib


def ref():
    """The reference's own letkf_core/mtx_eigen/distance_2p (FP32, K=128), or
    None where it was not built (the GPU host has no reference sources)."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_LIB):
            return None
        # MKL's default threading layer clashes with the libgomp that torch
        # loads into the same process (garbage results); the reference's
        # BLAS calls run inside its own OpenMP loop, i.e. sequentially.
        os.environ["MKL_THREADING_LAYER"] = "SEQUENTIAL"
        _ref = C.CDLL(REF_LIB)
        F = C.POINTER(C.c_float)
        D = C.POINTER(C.c_double)
        _ref.ref_nmem.restype = C.c_int
        _ref.ref_real_kind.restype = C.c_int
        _ref.ref_letkf_core.argtypes = [C.c_int, F, F, F, F, F, F]
        _ref.ref_letkf_core.restype = None
        _ref.ref_mtx_eigen.argtypes = [C.c_int, F, D, D]
        _ref.ref_mtx_eigen.restype = None
        _ref.ref_distance.argtypes = [C.c_float] * 4 + [F]
        _ref.ref_distance.restype = None
        _ref.ref_constants.argtypes = [F, F]
        _ref.ref_constants.restype = None
        I = C.POINTER(C.c_int32)
        _ref.ref_trans_xtoy.argtypes = [C.c_int, I, I, I, F, F, F, F, F, F, F, F, F, F]
        _ref.ref_trans_xtoy.restype = None
        _ref.ref_akima.argtypes = [C.c_int, F, F, C.c_float, F]
        _ref.ref_akima.restype = None
        _ref.ref_xtoy_consts.argtypes = [I, F]
        _ref.ref_xtoy_consts.restype = None
        _ref.ref_ensemble_mesp.argtypes = [C.c_int] + [F] * 6
        _ref.ref_ensemble_mesp.restype = None
    return _ref


def _suffix(dtype) -> str:
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def _ct(dtype):
    return C.c_double if np.dtype(dtype) == np.float64 else C.c_float


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _f(a, dtype):
    return np.asfortranarray(a, dtype=dtype)


# ---------------------------------------------------------------- kernels


def distance(lona, lata, lonb, latb, dtype=np.float64, pi=None, re=None):
    dt = np.dtype(dtype)
    pi = float(np.array(4.0 * np.arctan(1.0), dt)) if pi is None else pi
    re = float(np.array(6371.3e3, dt)) if re is None else re
    f = getattr(lib(), "oracle_distance" + _suffix(dt))
    return f(lona, lata, lonb, latb, pi, re)


def eigen(a, dtype=np.float64):
    a = _f(a, dtype)
    n = a.shape[0]
    ev = np.zeros(n)
    V = np.zeros((n, n), order="F")
    rc = getattr(lib(), "oracle_eigen" + _suffix(dtype))(
        n, _ptr(a, _ct(dtype)), _ptr(ev, C.c_double), _ptr(V, C.c_double))
    return rc, ev, V


def letkf_core(hdxf, rdiag, rloc, dep, cov_infl_mul=1.0, dtype=np.float64):
    ct = _ct(dtype)
    hdxf = _f(hdxf, dtype)
    n, K = hdxf.shape
    rdiag, rloc, dep = (_f(x, dtype) for x in (rdiag, rloc, dep))
    w = np.zeros(K, dtype)
    W = np.zeros((K, K), dtype, order="F")
    rc = getattr(lib(), "oracle_letkf_core" + _suffix(dtype))(
        n, K, cov_infl_mul, _ptr(hdxf, ct), _ptr(rdiag, ct), _ptr(rloc, ct),
        _ptr(dep, ct), _ptr(w, ct), _ptr(W, ct))
    return rc, w, W


@dataclass
class Problem:
    """One analysis in the reference's module-state form (host arrays,
    column-major, REAL = dtype).  Field names follow common_setting.f90."""
    params: dict
    lon1: np.ndarray
    lat1: np.ndarray
    ri1: np.ndarray
    rj1: np.ndarray
    dlon1: np.ndarray
    dlat1: np.ndarray
    phi1: np.ndarray
    depth1: np.ndarray          # (nij1, nlev)
    nobsgrd: np.ndarray         # (nlon, nlat) int32
    obselm: np.ndarray          # (nobs,) int32
    obslon: np.ndarray
    obslat: np.ndarray
    obslev: np.ndarray
    obserr: np.ndarray
    obsdep: np.ndarray
    obshdxf: np.ndarray         # (nobs, nmem)
    fcst3d: np.ndarray          # (nij1, nlev, nmem, nv3d)
    fcst2d: np.ndarray          # (nij1, nmem, nv2d)

    def astype(self, dtype) -> "Problem":
        kw = {}
        for k, v in self.__dict__.items():
            if isinstance(v, np.ndarray) and v.dtype.kind == "f":
                kw[k] = np.asfortranarray(v, dtype=dtype)
            else:
                kw[k] = v
        return Problem(**kw)


def make_params(pr: Problem, nthreads: int = 0) -> OracleParams:
    p = pr.params
    op = OracleParams()
    op.nmem = pr.obshdxf.shape[1]
    op.nlon, op.nlat = pr.nobsgrd.shape
    op.nij1, op.nlev = pr.depth1.shape
    op.nv3d = pr.fcst3d.shape[3]
    op.nv2d = pr.fcst2d.shape[2]
    for k in ("iv3d_t", "iv3d_s", "id_z_obs", "roff"):
        setattr(op, k, int(p[k]))
    op.nthreads = nthreads
    for k in ("pi", "re", "cov_infl_mul", "alpha_rtpp", "alpha_rtps",
              "sigma_obs", "sigma_obsv", "dist_zero", "dist_zerov", "tmin",
              "tmax", "smin", "smax"):
        setattr(op, k, float(p[k]))
    return op


def make_tables(pr: Problem, dtype):
    ct = _ct(dtype)
    T = Tables64 if np.dtype(dtype) == np.float64 else Tables32
    keep = []

    def P(a):
        a = _f(a, dtype)
        keep.append(a)
        return _ptr(a, ct)

    def I(a):
        a = np.asfortranarray(a, dtype=np.int32)
        keep.append(a)
        return a.ctypes.data_as(C.POINTER(C.c_int32))

    t = T(P(pr.lon1), P(pr.lat1), P(pr.ri1), P(pr.rj1), P(pr.dlon1), P(pr.dlat1),
          P(pr.phi1), P(pr.depth1), int(pr.obselm.shape[0]), I(pr.nobsgrd),
          I(pr.obselm), P(pr.obslon), P(pr.obslat), P(pr.obslev), P(pr.obserr),
          P(pr.obsdep), P(pr.obshdxf))
    t._keep = keep
    return t


def count_nobsl(pr: Problem, i: int, k: int, dtype=np.float64) -> int:
    op = make_params(pr)
    t = make_tables(pr, dtype)
    return getattr(lib(), "oracle_count_nobsl" + _suffix(dtype))(
        C.byref(op), C.byref(t), i, k)


def obs_local(pr: Problem, i: int, k: int, dtype=np.float64):
    ct = _ct(dtype)
    n = count_nobsl(pr, i, k, dtype)
    K = pr.obshdxf.shape[1]
    idx = np.zeros(max(n, 1), np.int32)
    hd = np.zeros((max(n, 1), K), dtype, order="F")
    rd, rl, dp = (np.zeros(max(n, 1), dtype) for _ in range(3))
    op = make_params(pr)
    t = make_tables(pr, dtype)
    got = getattr(lib(), "oracle_obs_local" + _suffix(dtype))(
        C.byref(op), C.byref(t), i, k, n, idx.ctypes.data_as(C.POINTER(C.c_int32)),
        _ptr(hd, ct), _ptr(rd, ct), _ptr(rl, ct), _ptr(dp, ct))
    assert got == n
    return idx[:n], hd[:n], rd[:n], rl[:n], dp[:n]


def da_main(pr: Problem, dtype=np.float64, nthreads: int = 0, lpfgm=None):
    """Full analysis.  Returns (rc, anal3d, anal2d, nobsl(nij1,nlev)).
    lpfgm: None (LETKF, iswitch_da 1) or a dict(mode=2|3, rand=(nmonte,
    nmem) draws, cov_infl_gm=1.0, rm="MR"|"SR", dp=True) selecting the
    LPF / LPFGM core (iswitch_da 2 / 3)."""
    ct = _ct(dtype)
    op = make_params(pr, nthreads)
    keep = None
    if lpfgm is not None:
        keep = np.ascontiguousarray(lpfgm["rand"], np.float64)
        op.iswitch_da = int(lpfgm.get("mode", 3))
        op.nmonte = keep.shape[0]
        op.rm_sr = 1 if lpfgm.get("rm", "MR") == "SR" else 0
        op.dp = 1 if lpfgm.get("dp", True) else 0
        op.cov_infl_gm = float(lpfgm.get("cov_infl_gm", 1.0))
        op.rand = keep.ctypes.data_as(C.POINTER(C.c_double))
    t = make_tables(pr, dtype)
    f3 = _f(pr.fcst3d, dtype)
    f2 = _f(pr.fcst2d, dtype)
    a3 = np.zeros_like(f3, order="F")
    a2 = np.zeros_like(f2, order="F")
    nl = np.zeros(pr.depth1.shape, np.int32, order="F")
    rc = getattr(lib(), "oracle_da_main" + _suffix(dtype))(
        C.byref(op), C.byref(t), _ptr(f3, ct), _ptr(f2, ct), _ptr(a3, ct),
        _ptr(a2, ct), nl.ctypes.data_as(C.POINTER(C.c_int32)))
    return rc, a3, a2, nl


def analyse_points(pr_tables: Problem, pt_i, pt_k, xf3, xf2, dtype=np.float64,
                   nthreads: int = 0, lpfgm=None):
    """Analysis of an explicit list of points.  xf3: (npt, nv3d, nmem) packed
    as the (nmem, nv3d) column-major slice per point; xf2 likewise.  lpfgm
    as for da_main."""
    ct = _ct(dtype)
    op = make_params(pr_tables, nthreads)
    keep = None
    if lpfgm is not None:
        keep = np.ascontiguousarray(lpfgm["rand"], np.float64)
        op.iswitch_da = int(lpfgm.get("mode", 3))
        op.nmonte = keep.shape[0]
        op.rm_sr = 1 if lpfgm.get("rm", "MR") == "SR" else 0
        op.dp = 1 if lpfgm.get("dp", True) else 0
        op.cov_infl_gm = float(lpfgm.get("cov_infl_gm", 1.0))
        op.rand = keep.ctypes.data_as(C.POINTER(C.c_double))
    t = make_tables(pr_tables, dtype)
    pt_i = np.ascontiguousarray(pt_i, np.int32)
    pt_k = np.ascontiguousarray(pt_k, np.int32)
    xf3 = np.ascontiguousarray(xf3, dtype)
    xf2 = np.ascontiguousarray(xf2, dtype)
    xa3 = np.zeros_like(xf3)
    xa2 = np.zeros_like(xf2)
    nl = np.zeros(len(pt_i), np.int32)
    IP = C.POINTER(C.c_int32)
    rc = getattr(lib(), "oracle_analyse_points" + _suffix(dtype))(
        C.byref(op), C.byref(t), len(pt_i), pt_i.ctypes.data_as(IP),
        pt_k.ctypes.data_as(IP), _ptr(xf3, ct), _ptr(xf2, ct), _ptr(xa3, ct),
        _ptr(xa2, ct), nl.ctypes.data_as(IP))
    return rc, xa3, xa2, nl


def aoei(err, dep, sprd, dtype=np.float64):
    ct = _ct(dtype)
    e = np.array(err, dtype)
    d = np.ascontiguousarray(dep, dtype)
    s = np.ascontiguousarray(sprd, dtype)
    getattr(lib(), "oracle_aoei" + _suffix(dtype))(len(e), _ptr(e, ct), _ptr(d, ct), _ptr(s, ct))
    return e


def bin_obs(lon2d, lat2d, olon, olat, dtype=np.float64):
    ct = _ct(dtype)
    lon2d = _f(lon2d, dtype)
    lat2d = _f(lat2d, dtype)
    nlon, nlat = lon2d.shape
    olon = np.ascontiguousarray(olon, dtype)
    olat = np.ascontiguousarray(olat, dtype)
    n = len(olon)
    IP = C.POINTER(C.c_int32)
    idx = np.zeros(n, np.int32)
    idy = np.zeros(n, np.int32)
    perm = np.zeros(max(n, 1), np.int32)
    grd = np.zeros((nlon, nlat), np.int32, order="F")
    s = getattr(lib(), "oracle_bin_obs" + _suffix(dtype))(
        nlon, nlat, _ptr(lon2d, ct), _ptr(lat2d, ct), n, _ptr(olon, ct),
        _ptr(olat, ct), idx.ctypes.data_as(IP), idy.ctypes.data_as(IP),
        perm.ctypes.data_as(IP), grd.ctypes.data_as(IP))
    return idx, idy, perm[:s], grd


# ---------------------------------------------------------- H(x) operator


def akima(x, y, x5, undef=999.0, dtype=np.float64):
    """akima_spline (common.f90:276-385) at one target."""
    dt = np.dtype(dtype)
    ct = _ct(dt)
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    f = getattr(lib(), "oracle_akima" + _suffix(dt))
    return float(f(len(x), _ptr(x, ct), _ptr(y, ct), ct(float(x5)), ct(float(undef))))


def trans_xtoy(g: OracleXtoY, elm, idx, idy, lono, lato, levo, lon, lat, depth, fsm, v3d, v2d,
               dtype=np.float64):
    """Trans_XtoY (common.f90:391-500) of one state; arrays in the Fortran
    layouts (numpy arrays indexed [i, j, ...] in column-major order)."""
    dt = np.dtype(dtype)
    ct = _ct(dt)
    n = len(elm)
    I = lambda a: np.ascontiguousarray(a, np.int32)
    elm, idx, idy = I(elm), I(idx), I(idy)
    lono, lato, levo = (np.ascontiguousarray(a, dt) for a in (lono, lato, levo))
    lon, lat, depth, fsm, v3d, v2d = (_f(a, dt) for a in (lon, lat, depth, fsm, v3d, v2d))
    hx = np.empty(n, dt)
    IP = C.POINTER(C.c_int32)
    getattr(lib(), "oracle_trans_xtoy" + _suffix(dt))(
        C.byref(g), n, elm.ctypes.data_as(IP), idx.ctypes.data_as(IP), idy.ctypes.data_as(IP),
        _ptr(lono, ct), _ptr(lato, ct), _ptr(levo, ct), _ptr(lon, ct), _ptr(lat, ct), _ptr(depth, ct),
        _ptr(fsm, ct), _ptr(v3d, ct), _ptr(v2d, ct), _ptr(hx, ct))
    return hx



def set_obs(c: OracleSetObs, elm, lon, lat, lev, dat, err, idx, idy, hx, qc=None, ins=None,
            nobslots=None, dtype=np.float64):
    """set_obs after H(x) (common_obs.f90:139-402); hx is (nobs, nmem).
    Returns a dict of the sorted tables (hdxf (nobs_out, nmem)), nobsgrd
    (nlon, nlat) and ndel = (numdelqc1, numdelqc2).  A nobslots sequence
    sets c.nslots to its length."""
    dt = np.dtype(dtype)
    ct = _ct(dt)
    n = len(elm)
    IP = C.POINTER(C.c_int32)
    I = lambda a: np.ascontiguousarray(a, np.int32)
    R = lambda a: np.ascontiguousarray(a, dt)
    elm, idx, idy = I(elm), I(idx), I(idy)
    lon, lat, lev, dat, err = (R(a) for a in (lon, lat, lev, dat, err))
    hxf = np.asfortranarray(np.asarray(hx, dt).reshape(n, c.nmem))
    qca = None if qc is None else I(qc)
    insa = None if ins is None else I(ins)
    sl = None if nobslots is None else I(nobslots)
    if sl is not None:
        c.nslots = len(sl)
    out = {k: np.zeros(max(n, 1), np.int32) for k in ("elm", "ins", "idx", "idy")}
    out.update({k: np.zeros(max(n, 1), dt) for k in ("lon", "lat", "lev", "dat", "err", "dep",
                                                      "hxfmean", "hxfsprd")})
    hd = np.zeros(max(n, 1) * c.nmem, dt)
    grd = np.zeros(c.nlon * c.nlat, np.int32)
    ndel = np.zeros(2, np.int32)
    ip = lambda a: None if a is None else a.ctypes.data_as(IP)
    nout = getattr(lib(), "oracle_set_obs" + _suffix(dt))(
        C.byref(c), n, ip(sl), ip(elm), ip(insa), _ptr(lon, ct), _ptr(lat, ct), _ptr(lev, ct),
        _ptr(dat, ct), _ptr(err, ct), ip(idx), ip(idy), _ptr(hxf, ct), ip(qca),
        ip(out["elm"]), ip(out["ins"]), ip(out["idx"]), ip(out["idy"]),
        *(_ptr(out[k], ct) for k in ("lon", "lat", "lev", "dat", "err", "dep", "hxfmean", "hxfsprd")),
        _ptr(hd, ct), ip(grd), ip(ndel))
    if nout < 0:
        raise ValueError("nobslots does not sum to nobs")
    res = {k: v[:nout].copy() for k, v in out.items()}
    res["hdxf"] = hd[:nout * c.nmem].reshape(c.nmem, nout).T.copy()
    res["nobsgrd"] = grd.reshape(c.nlat, c.nlon).T.copy()
    res["nobs"] = nout
    res["ndel"] = (int(ndel[0]), int(ndel[1]))
    return res


def ensemble_mesp(v3d, v2d, undef=999.0, dtype=np.float64):
    """ensemble_mesp (common.f90:505-596): v3d (nij1,nlev,nmem,nv3d), v2d
    (nij1,nmem,nv2d) in the Fortran layouts; returns v3dm, v3ds
    (nij1,nlev,nv3d) and v2dm, v2ds (nij1,nv2d)."""
    ct = _ct(dtype)
    a3, a2 = _f(v3d, dtype), _f(v2d, dtype)
    nij1, nlev, nmem, nv3d = a3.shape
    nv2d = a2.shape[2]
    if a2.shape[:2] != (nij1, nmem):
        raise ValueError("v2d must be (nij1, nmem, nv2d)")
    out = [np.zeros((nij1, nlev, nv3d), dtype, order="F") for _ in range(2)]
    out += [np.zeros((nij1, nv2d), dtype, order="F") for _ in range(2)]
    getattr(lib(), "oracle_ensemble_mesp" + _suffix(dtype))(nij1, nlev, nmem, nv3d, nv2d, undef, _ptr(a3, ct),
                                                          _ptr(a2, ct), *(_ptr(o, ct) for o in out))
    return tuple(out)


def grid_pack(nlon, nlat, v3d, v2d, wrap=True, file_order=False, dtype=np.float64):
    """Single-precision global grids of several fields as the output step
    writes them (gather_grd_mpi on one rank, common_mpi.f90:133-206; the
    quasi-global wrap, common_io.f90:960-963; the file's vertical order,
    common_io.f90:655-661).  v3d (nij1,nlev,nfield,nv3d), v2d
    (nij1,nfield,nv2d) in the Fortran layouts of the analysis arrays (the
    fields being members); returns g3 (nlon,nlat,nlev,nv3d,nfield) and g2
    (nlon,nlat,nv2d,nfield), Fortran order."""
    ct = _ct(dtype)
    a3, a2 = _f(v3d, dtype), _f(v2d, dtype)
    nij, nlev, nf, nv3d = a3.shape
    nv2d = a2.shape[2]
    if nij != nlon * nlat or a2.shape[:2] != (nij, nf):
        raise ValueError("v3d (nlon*nlat, nlev, nfield, nv3d), v2d (nlon*nlat, nfield, nv2d)")
    g3 = np.zeros((nlon, nlat, nlev, nv3d, nf), np.float32, order="F")
    g2 = np.zeros((nlon, nlat, nv2d, nf), np.float32, order="F")
    getattr(lib(), "oracle_grid_pack" + _suffix(dtype))(
        nlon, nlat, nlev, nv3d, nv2d, nf, int(wrap), int(file_order), _ptr(a3, ct), nij * nlev, nij * nlev * nf,
        nij, _ptr(a2, ct), nij, nij * nf, _ptr(g3, C.c_float), _ptr(g2, C.c_float))
    return g3, g2


def lpfgm_core(hdxf, rdiag, rloc, dep, rand, mode=3, cov_infl_gm=1.0, rm="MR", dp=True, dtype=np.float64):
    """lpfgm_core (common_lpfgm.f90:444-494): mode 2 = LPF, 3 = LPFGM.
    hdxf (nobsl, nmem); rand (nmonte, nmem) uniforms in [0, 1), the draws of
    the nmonte resamplings.  Returns rc, wvec (nmem), Wmat (nmem, nmem)."""
    ct = _ct(dtype)
    hdxf = _f(hdxf, dtype)
    n, K = hdxf.shape
    rdiag, rloc, dep = (_f(x, dtype) for x in (rdiag, rloc, dep))
    r = np.ascontiguousarray(rand, np.float64)
    if r.ndim != 2 or r.shape[1] != K:
        raise ValueError("rand must be (nmonte, nmem)")
    w = np.zeros(K, dtype)
    W = np.zeros((K, K), dtype, order="F")
    rc = getattr(lib(), "oracle_lpfgm_core" + _suffix(dtype))(
        n, K, mode, cov_infl_gm, r.shape[0], 1 if rm == "SR" else 0, 1 if dp else 0,
        r.ctypes.data_as(C.POINTER(C.c_double)), _ptr(hdxf, ct), _ptr(rdiag, ct), _ptr(rloc, ct), _ptr(dep, ct),
        _ptr(w, ct), _ptr(W, ct))
    return rc, w, W

