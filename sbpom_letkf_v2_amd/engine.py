"""Host-side mirror of the reference's analysis entry point.

`da_main(problem)` plays the role of
    CALL da_main(fcst3d, fcst2d, anal3d, anal2d)      (letkf/run/src/main.f90:92)
with the module state of common_setting.f90 passed explicitly as a
`synth.Problem`-like object.  Host numpy arrays go through the engine's
host entry point (letkf_da_main), torch tensors on the GPU through the
device entry point (letkf_da_main_device).  Errors raise EngineError with the
engine's status code instead of STOP.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (FORM_BATCHED, FORM_BATCHED_QL, FORM_PLAIN, PREC_BF16X3, PREC_BF16X3_TRI32, PREC_FP32,
                   PREC_FP32_DUAL64, PREC_FP32_PANEL, PREC_FP32X3, PREC_FP64, check)

PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "fp32p": PREC_FP32_PANEL,
              "fp32x3": PREC_FP32X3, "fp32d": PREC_FP32_DUAL64, "bf16x3t": PREC_BF16X3_TRI32}
FORMS = {"plain": FORM_PLAIN, "batched": FORM_BATCHED, "batched_ql": FORM_BATCHED_QL}


def _shape_of(a):
    """Fortran shape of a numpy array or of a tensor holding a Fortran layout."""
    if isinstance(a, np.ndarray):
        return a.shape
    return tuple(reversed(a.shape))


def _is_torch(a):
    return type(a).__module__.startswith("torch")


def make_params(pr, precision="fp64", form="plain") -> _lib.Params:
    p = pr.params
    nij1, nlev = _shape_of(pr.depth1)
    nlon, nlat = _shape_of(pr.nobsgrd)
    f3 = _shape_of(pr.fcst3d)
    f2 = _shape_of(pr.fcst2d)
    prm = _lib.Params()
    check(_lib.lib().letkf_params_default(C.byref(prm), int(p.get("real_kind", 8))))
    prm.nmem = f3[2]
    prm.nlon, prm.nlat, prm.nlev, prm.nij1 = nlon, nlat, nlev, nij1
    prm.nv3d = f3[3]
    prm.nv2d = f2[2] if len(f2) == 3 else 0
    for k in ("iv3d_t", "iv3d_s", "id_z_obs", "roff"):
        if k in p:
            setattr(prm, k, int(p[k]))
    for k in ("pi", "re", "cov_infl_mul", "alpha_rtpp", "alpha_rtps", "sigma_obs",
              "sigma_obsv", "dist_zero", "dist_zerov", "tmin", "tmax", "smin", "smax"):
        if k in p:
            setattr(prm, k, float(p[k]))
    prm.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    prm.form = FORMS[form] if isinstance(form, str) else int(form)
    return prm


class Engine:
    """One engine handle (device workspaces + a HIP stream)."""

    def __init__(self, params: _lib.Params):
        self.params = params
        h = C.c_void_p()
        check(_lib.lib().letkf_create(C.byref(params), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _lib.lib().letkf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def stats(self) -> dict:
        s = _lib.Stats()
        check(_lib.lib().letkf_get_stats(self.h, C.byref(s)))
        out = {}
        for n, _ in s._fields_:
            v = getattr(s, n)
            out[n] = list(v) if n in ("ms_eigk", "n_eigk", "bucket_slots") else v
        return out

    def nobsl(self):
        """letkf_get_nobsl: per-point local-obs counts of the last analysis
        or plan, an int32 device tensor (nlev, nij1) (memory = nobsl(nij1,nlev))."""
        import torch
        t = torch.empty((int(self.params.nlev), int(self.params.nij1)), dtype=torch.int32, device="cuda")
        check(_lib.lib().letkf_get_nobsl(self.h, _ptr(t), None))
        return t

    # -------------------------------------------------------- host arrays
    def da_main_host(self, pr, lpfgm=None):
        """letkf_da_main on host arrays (Fortran layouts), or with lpfgm =
        dict(mode, rand, cov_infl_gm, rm, dp) letkf_da_main_lpfgm."""
        dt = np.float32 if self.params.real_kind == 4 else np.float64
        keep = []

        def P(a, dtype=dt):
            a = np.asfortranarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        g = _lib.Grid(*(P(getattr(pr, n)) for n in (
            "lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1", "depth1")))
        o = _lib.Obs(int(pr.obselm.shape[0]), P(pr.nobsgrd, np.int32), P(pr.obselm, np.int32),
                     *(P(getattr(pr, n)) for n in ("obslon", "obslat", "obslev", "obserr",
                                                    "obsdep", "obshdxf")))
        f3 = np.asfortranarray(pr.fcst3d, dtype=dt)
        f2 = np.asfortranarray(pr.fcst2d, dtype=dt)
        a3 = np.zeros_like(f3, order="F")
        a2 = np.zeros_like(f2, order="F")
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        if lpfgm is None:
            check(_lib.lib().letkf_da_main(self.h, C.byref(g), C.byref(o), vp(f3), vp(f2), vp(a3), vp(a2)))
            return a3, a2
        r = np.ascontiguousarray(lpfgm["rand"], np.float64)
        cfg = _lib.LpfgmCfg(int(lpfgm.get("mode", 3)), r.shape[0], 1 if lpfgm.get("rm", "MR") == "SR" else 0,
                            1 if lpfgm.get("dp", True) else 0, float(lpfgm.get("cov_infl_gm", 1.0)),
                            r.ctypes.data_as(C.POINTER(C.c_double)))
        check(_lib.lib().letkf_da_main_lpfgm(self.h, C.byref(g), C.byref(o), C.byref(cfg), vp(f3), vp(f2), vp(a3),
                                             vp(a2)))
        return a3, a2

    def set_params(self, params: _lib.Params):
        """letkf_set_params: new scalars / precision / form, same sizes."""
        check(_lib.lib().letkf_set_params(self.h, C.byref(params)))
        self.params = params

    # ------------------------------------------------------ device arrays
    def _device_structs(self, pr):
        import torch
        dt = torch.float32 if self.params.real_kind == 4 else torch.float64
        p = self.params
        nij1, nlev, K = int(p.nij1), int(p.nlev), int(p.nmem)
        nobs = int(pr.obselm.numel())
        # every array the engine reads: (name, dtype, element count)
        need = [(n, dt, nij1) for n in ("lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1")]
        need += [("depth1", dt, nij1 * nlev), ("nobsgrd", torch.int32, int(p.nlon) * int(p.nlat)),
                 ("obselm", torch.int32, nobs)]
        need += [(n, dt, nobs) for n in ("obslon", "obslat", "obslev", "obserr", "obsdep")]
        need += [("obshdxf", dt, nobs * K), ("fcst3d", dt, nij1 * nlev * K * int(p.nv3d)),
                 ("fcst2d", dt, nij1 * K * int(p.nv2d))]
        for n, t, cnt in need:
            a = getattr(pr, n)
            if not (_is_torch(a) and a.is_cuda):
                raise TypeError(f"{n} must be a torch tensor on the GPU")
            if a.dtype != t:
                raise TypeError(f"{n} has dtype {a.dtype}, needs {t}")
            if not a.is_contiguous():
                raise ValueError(f"{n} must be contiguous")
            if a.numel() != cnt:
                raise ValueError(f"{n} has {a.numel()} elements, the handle's sizes need {cnt}")
        g = _lib.Grid(*(_ptr(getattr(pr, n)) for n in (
            "lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1", "depth1")))
        o = _lib.Obs(int(pr.obselm.numel()), _ptr(pr.nobsgrd), _ptr(pr.obselm),
                     *(_ptr(getattr(pr, n)) for n in ("obslon", "obslat", "obslev", "obserr",
                                                      "obsdep", "obshdxf")))
        return g, o

    def da_main_device(self, pr, anal3d=None, anal2d=None, stream=None):
        import torch
        g, o = self._device_structs(pr)
        ptr = _ptr
        if anal3d is None:
            anal3d = torch.empty_like(pr.fcst3d)
        if anal2d is None:
            anal2d = torch.empty_like(pr.fcst2d)
        s = C.c_void_p(stream) if stream else None
        check(_lib.lib().letkf_da_main_device(self.h, C.byref(g), C.byref(o), ptr(pr.fcst3d),
                                              ptr(pr.fcst2d), ptr(anal3d), ptr(anal2d), s))
        return anal3d, anal2d

    def da_main_lpfgm_device(self, pr, rand, mode=3, cov_infl_gm=1.0, rm="MR", dp=True, anal3d=None,
                             anal2d=None, stream=None):
        """letkf_da_main_lpfgm_device: da_main with iswitch_da = mode (2 LPF,
        3 LPFGM; common_da.f90:74-84, common_lpfgm.f90).  rand (nmonte, nmem)
        uniforms in [0, 1) are the draws of the resamplings."""
        import numpy as np
        import torch
        g, o = self._device_structs(pr)
        r = np.ascontiguousarray(rand, np.float64)
        if r.ndim != 2 or r.shape[1] != self.params.nmem:
            raise ValueError("rand must be (nmonte, nmem)")
        cfg = _lib.LpfgmCfg(int(mode), r.shape[0], 1 if rm == "SR" else 0, 1 if dp else 0, float(cov_infl_gm),
                            r.ctypes.data_as(C.POINTER(C.c_double)))
        if anal3d is None:
            anal3d = torch.empty_like(pr.fcst3d)
        if anal2d is None:
            anal2d = torch.empty_like(pr.fcst2d)
        s = C.c_void_p(stream) if stream else None
        check(_lib.lib().letkf_da_main_lpfgm_device(self.h, C.byref(g), C.byref(o), C.byref(cfg), _ptr(pr.fcst3d),
                                                    _ptr(pr.fcst2d), _ptr(anal3d), _ptr(anal2d), s))
        return anal3d, anal2d

    BUCKETS = (16, 32, 48, 64, 128)

    def autotune(self, pr, anal3d=None, anal2d=None, stream=None):
        """letkf_autotune: time the vector paths of every size bucket on the
        first chunk of pr's analysis, keep the fastest, run the analysis.
        Returns (anal3d, anal2d, {bucket: {path: ms}} for the paths
        measured (1 Q, 2 reflectors, 3 blocked reflectors), chosen paths)."""
        import torch
        g, o = self._device_structs(pr)
        if anal3d is None:
            anal3d = torch.empty_like(pr.fcst3d)
        if anal2d is None:
            anal2d = torch.empty_like(pr.fcst2d)
        ms = (C.c_double * 15)()
        check(_lib.lib().letkf_autotune(self.h, C.byref(g), C.byref(o), _ptr(pr.fcst3d), _ptr(pr.fcst2d),
                                        _ptr(anal3d), _ptr(anal2d), ms, C.c_void_p(stream) if stream else None))
        times = {}
        for q, nb in enumerate(self.BUCKETS):
            t = {v: ms[3 * q + v - 1] for v in (1, 2, 3) if ms[3 * q + v - 1] >= 0}
            if t:
                times[nb] = t
        return anal3d, anal2d, times, self.get_tuning()

    def get_tuning(self):
        v = (C.c_int32 * 5)()
        check(_lib.lib().letkf_get_tuning(self.h, v))
        return dict(zip(self.BUCKETS, list(v)))

    def set_tuning(self, paths):
        """paths: {bucket: 0 default rule | 1 Q | 2 reflectors | 3 blocked
        reflectors} or 5 values."""
        vals = [paths.get(nb, 0) for nb in self.BUCKETS] if isinstance(paths, dict) else list(paths)
        check(_lib.lib().letkf_set_tuning(self.h, (C.c_int32 * 5)(*vals)))

    def plan_device(self, pr, stream=None):
        """letkf_da_main_plan: the geometry-dependent part of the analysis of
        pr (local-obs counts, chunk plan, size buckets, workspaces)."""
        g, o = self._device_structs(pr)
        check(_lib.lib().letkf_da_main_plan(self.h, C.byref(g), C.byref(o), C.c_void_p(stream) if stream else None))

    def enqueue_device(self, pr, anal3d, anal2d, stream=None):
        """letkf_da_main_enqueue: the analysis of pr on the current plan as
        kernel launches on `stream` only (capturable in a graph).  pr's
        arrays must be the planned ones; their values may change."""
        g, o = self._device_structs(pr)
        check(_lib.lib().letkf_da_main_enqueue(self.h, C.byref(g), C.byref(o), _ptr(pr.fcst3d), _ptr(pr.fcst2d),
                                               _ptr(anal3d), _ptr(anal2d), C.c_void_p(stream) if stream else None))

    def status(self, stream=None):
        """letkf_da_main_status: wait for `stream`, raise on kernel status bits."""
        check(_lib.lib().letkf_da_main_status(self.h, C.c_void_p(stream) if stream else None))

    def set_obs_device(self, nlon, nlat, elm, lon, lat, lev, dat, err, idx, idy, hx, qc=None, ins=None,
                       nobslots=None, **cfg):
        """letkf_set_obs_device: set_obs after H(x) (common_obs.f90:139-402)
        on device tensors of the nobs unsorted obs; hx (nmem, nobs) and qc as
        trans_xtoy_device returns them.  cfg overrides the reference's
        settings (lge_io, aoei, nslots, nbslot, sigma_obst, undef,
        gross_error, id_*_obs, qc_min / qc_max as 9-sequences); nobslots is
        a host sequence of per-slot counts.  Returns a dict of device tensors
        named as the obs tables da_main_device reads (obselm, obslon, obslat,
        obslev, obserr, obsdep, obshdxf (nmem, nobs_out), nobsgrd (nlat,
        nlon)) plus obsins, obsidx, obsidy, obsdat, obshxfmean, obshxfsprd,
        and the counts nobs, ndel_proj, ndel_gross."""
        import torch
        dt = torch.float32 if self.params.real_kind == 4 else torch.float64
        nmem = int(self.params.nmem)
        nobs = int(elm.numel())
        for a in (lon, lat, lev, dat, err, hx):
            if a.dtype != dt or not a.is_contiguous() or not a.is_cuda:
                raise ValueError(f"real arrays must be contiguous {dt} device tensors")
        for a in (elm, idx, idy) + tuple(x for x in (qc, ins) if x is not None):
            if a.dtype != torch.int32 or not a.is_contiguous() or not a.is_cuda:
                raise ValueError("integer arrays must be contiguous int32 device tensors")
        if hx.numel() != nobs * nmem:
            raise ValueError(f"hx must hold nmem*nobs = {nmem}*{nobs} values")
        c = _lib.SetObsCfg()
        _lib.lib().letkf_setobs_default(C.byref(c))
        for k, v in cfg.items():
            if k in ("qc_min", "qc_max"):
                for i, x in enumerate(v):
                    getattr(c, k)[i] = float(x)
            elif hasattr(c, k):
                setattr(c, k, v)
            else:
                raise TypeError(f"unknown set_obs setting {k}")
        slots = None
        if nobslots is not None:
            nobslots = [int(x) for x in nobslots]
            c.nslots = len(nobslots)
            slots = (C.c_int32 * len(nobslots))(*nobslots)
        dev = elm.device
        cap = max(nobs, 1)
        I = lambda: torch.empty(cap, dtype=torch.int32, device=dev)
        R = lambda: torch.empty(cap, dtype=dt, device=dev)
        o = {"obselm": I(), "obsins": I(), "obsidx": I(), "obsidy": I(), "obslon": R(), "obslat": R(),
             "obslev": R(), "obsdat": R(), "obserr": R(), "obsdep": R(), "obshxfmean": R(), "obshxfsprd": R()}
        hd = torch.empty(cap * nmem, dtype=dt, device=dev)
        grd = torch.empty((nlat, nlon), dtype=torch.int32, device=dev)
        out = _lib.ObsSet(_ptr(grd), *(_ptr(o[k]) for k in ("obselm", "obsins", "obsidx", "obsidy")),
                          *(_ptr(o[k]) for k in ("obslon", "obslat", "obslev", "obsdat", "obserr", "obsdep",
                                                 "obshxfmean", "obshxfsprd")), _ptr(hd), 0, 0, 0)
        check(_lib.lib().letkf_set_obs_device(self.h, C.byref(c), nlon, nlat, nobs, slots, _ptr(elm), _ptr(ins),
                                              _ptr(lon), _ptr(lat), _ptr(lev), _ptr(dat), _ptr(err), _ptr(idx),
                                              _ptr(idy), _ptr(hx), _ptr(qc), C.byref(out)))
        n = int(out.nobs_out)
        res = {k: v[:n] for k, v in o.items()}
        res["obshdxf"] = hd[:n * nmem].view(nmem, n)
        res["nobsgrd"] = grd
        res.update(nobs=n, ndel_proj=int(out.ndel_proj), ndel_gross=int(out.ndel_gross))
        return res

    def da_adjoint_device(self, pr, g3, g2, stream=None):
        """letkf_da_adjoint_device: (dL/d obsdep, dL/d log obserr^2,
        [dL/d alpha_rtpp, dL/d cov_infl_mul, dL/d sigma_obs, dL/d sigma_obsv])
        for upstream gradients g3, g2 of (anal3d, anal2d); pr and the
        handle's parameters as in the forward call."""
        import torch
        g, o = self._device_structs(pr)
        dev = pr.fcst3d.device
        nobs = int(pr.obselm.numel())
        gd = torch.zeros(nobs, dtype=torch.float64, device=dev)
        gl = torch.zeros(nobs, dtype=torch.float64, device=dev)
        gs = torch.zeros(4, dtype=torch.float64, device=dev)
        g3 = g3.contiguous()
        g2 = g2.contiguous() if g2 is not None else None
        s = C.c_void_p(stream) if stream else None
        check(_lib.lib().letkf_da_adjoint_device(self.h, C.byref(g), C.byref(o), _ptr(pr.fcst3d),
                                                 _ptr(pr.fcst2d), _ptr(g3), _ptr(g2), _ptr(gd), _ptr(gl),
                                                 _ptr(gs), s))
        return gd, gl, gs


def _ptr(a):
    return C.c_void_p(a.data_ptr()) if a is not None and a.numel() else None


def trans_xtoy_device(lon, lat, depth, fsm, elm, idx, idy, lono, lato, levo, v3d, v2d, qc=True,
                      id_obs=(2819, 2820, 3073, 3332, 2567), iv3d=(1, 2, 3, 4), iv2d_z=1, hnosea=0.0,
                      undef=999.0, analysis_layout=False):
    """H(x) of every member (letkf_trans_xtoy_device; Trans_XtoY,
    letkf/run/src/common.f90:391-500, batched over members).

    Device tensors in the reference's Fortran layouts, held as C-contiguous
    torch tensors with the axes reversed: lon, lat, fsm (nlat, nlon); depth
    (nlev, nlat, nlon); v3d (nmem, nv3d, nlev, nlat, nlon); v2d (nmem, nv2d,
    nlat, nlon); obs vectors (nobs).  With analysis_layout the members are
    the analysis arrays instead: v3d (nv3d, nmem, nlev, nlat*nlon) =
    fcst3d(nij1,nlev,nmem,nv3d), v2d (nv2d, nmem, nlat*nlon) =
    fcst2d(nij1,nmem,nv2d).  Returns hx (nmem, nobs), i.e. hx(nobs, nmem) of
    the reference, and qc (nobs, int32) or None.  id_obs are the element ids
    of u, v, t, s, SSH; iv3d the 1-based slots of u, v, t, s."""
    import torch
    dt = v3d.dtype
    strides = None
    if analysis_layout:
        nv3d, nmem, nlev, nij = v3d.shape
        nv2d = v2d.shape[0]
        nlat, nlon = lon.shape
        if nij != nlat * nlon or v2d.shape[1] != nmem:
            raise ValueError("analysis layout: v3d (nv3d, nmem, nlev, nlat*nlon), v2d (nv2d, nmem, nlat*nlon)")
        strides = (C.c_int64 * 4)(nij * nlev * nmem, nij * nlev, nij * nmem, nij)
    else:
        nmem, nv3d, nlev, nlat, nlon = v3d.shape
        nv2d = v2d.shape[1]
    for a in (lon, lat, depth, fsm, lono, lato, levo, v2d):
        if a.dtype != dt or not a.is_contiguous() or not a.is_cuda:
            raise ValueError("real arrays must be contiguous device tensors of one dtype")
    for a in (elm, idx, idy):
        if a.dtype != torch.int32 or not a.is_contiguous() or not a.is_cuda:
            raise ValueError("elm/idx/idy must be contiguous int32 device tensors")
    nobs = int(elm.numel())
    g = _lib.XtoyGrid(nlon, nlat, nlev, nv3d, nv2d, 8 if dt == torch.float64 else 4, *id_obs, *iv3d, iv2d_z,
                      float(hnosea), float(undef), _ptr(lon), _ptr(lat), _ptr(depth), _ptr(fsm))
    hx = torch.empty((nmem, nobs), dtype=dt, device=v3d.device)
    q = torch.empty(nobs, dtype=torch.int32, device=v3d.device) if qc else None
    s = C.c_void_p(torch.cuda.current_stream(v3d.device).cuda_stream)
    _lib.check(_lib.lib().letkf_trans_xtoy_device(C.byref(g), nobs, _ptr(elm), _ptr(idx), _ptr(idy),
                                                   _ptr(lono), _ptr(lato), _ptr(levo), nmem, _ptr(v3d),
                                                   _ptr(v2d), strides, _ptr(hx), _ptr(q), s))
    return hx, q


def ensemble_mesp_device(v3d, v2d, undef=999.0):
    """Grid ensemble mean and spread (letkf_ensemble_mesp_device;
    ensemble_mesp, letkf/run/src/common.f90:505-596) of device tensors in the
    analysis layout: v3d (nv3d, nmem, nlev, nij1) = v3d(nij1,nlev,nmem,nv3d),
    v2d (nv2d, nmem, nij1).  Returns v3dm, v3ds (nv3d, nlev, nij1) and v2dm,
    v2ds (nv2d, nij1)."""
    import torch
    nv3d, nmem, nlev, nij = v3d.shape
    nv2d = v2d.shape[0]
    dt = v3d.dtype
    if v2d.shape[1:] != (nmem, nij) or v2d.dtype != dt or dt not in (torch.float32, torch.float64):
        raise ValueError("v3d (nv3d, nmem, nlev, nij1), v2d (nv2d, nmem, nij1) of one real dtype")
    for a in (v3d, v2d):
        if not (a.is_cuda and a.is_contiguous()):
            raise ValueError("v3d and v2d must be contiguous device tensors")
    m3 = torch.empty((nv3d, nlev, nij), dtype=dt, device=v3d.device)
    s3 = torch.empty_like(m3)
    m2 = torch.empty((nv2d, nij), dtype=dt, device=v3d.device)
    s2 = torch.empty_like(m2)
    st = C.c_void_p(torch.cuda.current_stream(v3d.device).cuda_stream)
    _lib.check(_lib.lib().letkf_ensemble_mesp_device(8 if dt == torch.float64 else 4, nij, nlev, nmem, nv3d,
                                                      nv2d, float(undef), _ptr(v3d), _ptr(v2d), _ptr(m3),
                                                      _ptr(s3), _ptr(m2), _ptr(s2), st))
    return m3, s3, m2, s2


def grid_pack_device(nlon, nlat, v3d, v2d, wrap=True, file_order=False, g3=None, g2=None):
    """REAL(4) global grids of every field of v3d/v2d (letkf_grid_pack_device:
    gather_grd_mpi on one rank, the quasi-global wrap, and with file_order
    the file's vertical order; letkf/run/src/common_io.f90:863-1074).
    v3d (nv3d, nfield, nlev, nlat*nlon) and v2d (nv2d, nfield, nlat*nlon) in
    the analysis layout (fields = members), or (nv3d, nlev, nij1) / (nv2d,
    nij1) for one field (a mean or a spread).  Returns g3 (nfield, nv3d,
    nlev, nlat, nlon) and g2 (nfield, nv2d, nlat, nlon) float32, i.e.
    v3dg(nlon,nlat,nlev,nv3d) / v2dg(nlon,nlat,nv2d) per field."""
    import torch
    one = v3d.dim() == 3
    if one:
        v3d, v2d = v3d.unsqueeze(1), v2d.unsqueeze(1)
    nv3d, nf, nlev, nij = v3d.shape
    nv2d = v2d.shape[0]
    dt = v3d.dtype
    if nij != nlon * nlat or v2d.shape[1:] != (nf, nij) or v2d.dtype != dt:
        raise ValueError("v3d (nv3d, nfield, nlev, nlon*nlat), v2d (nv2d, nfield, nlon*nlat) of one dtype")
    for a in (v3d, v2d):
        if not (a.is_cuda and a.is_contiguous()):
            raise ValueError("v3d and v2d must be contiguous device tensors")
    if g3 is None:
        g3 = torch.empty((nf, nv3d, nlev, nlat, nlon), dtype=torch.float32, device=v3d.device)
    if g2 is None:
        g2 = torch.empty((nf, nv2d, nlat, nlon), dtype=torch.float32, device=v3d.device)
    if g3.shape != (nf, nv3d, nlev, nlat, nlon) or g2.shape != (nf, nv2d, nlat, nlon):
        raise ValueError("g3/g2 shapes")
    s3 = (C.c_int64 * 3)(nij * nlev, nij * nlev * nf, nij)
    s2 = (C.c_int64 * 2)(nij, nij * nf)
    st = C.c_void_p(torch.cuda.current_stream(v3d.device).cuda_stream)
    _lib.check(_lib.lib().letkf_grid_pack_device(8 if dt == torch.float64 else 4, nlon, nlat, nlev, nv3d, nv2d,
                                                  nf, int(wrap), int(file_order), _ptr(v3d), s3, _ptr(v2d), s2,
                                                  _ptr(g3), _ptr(g2), st))
    return g3, g2


def da_main(pr, precision="fp64", form="plain"):
    """Analysis of one problem; returns (anal3d, anal2d) in the same kind
    of container as the inputs."""
    eng = Engine(make_params(pr, precision, form))
    try:
        if _is_torch(pr.fcst3d):
            return eng.da_main_device(pr)
        return eng.da_main_host(pr)
    finally:
        eng.close()
