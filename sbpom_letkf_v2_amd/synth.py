"""Synthetic analysis problems in the reference's module-state form.

A problem is what set_obs (letkf/run/src/common_obs.f90:24-402) leaves in
common_setting for da_main: the local grid (lon1, lat1, ri1, rj1, dlon1,
dlat1, phi1, depth1), the sorted observation tables (nobsgrd, obselm,
obslon, obslat, obslev, obserr, obsdep, obshdxf) and the forecast ensemble
fcst3d(nij1,nlev,nmem,nv3d), fcst2d(nij1,nmem,nv2d), for one process
holding the whole grid (nij1 = nlon*nlat, i fastest).

Fields are Gaussian random fields: white noise smoothed by a separable
Gaussian kernel of std L/2 cells (correlation exp(-r^2/L^2)), normalised to
unit variance.  Observations sit on grid nodes, so H is point sampling and
the obs cell (obs_id_uniform, common.f90:112-168) is the node's own cell.
Every real value is rounded to the requested REAL before use, so FP32 and
FP64 consumers see identical inputs.

Generation runs in torch on any device: the CPU for test sizes, the GPU for
the 512x512x40 workloads.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

ID_U, ID_V, ID_T, ID_S, ID_Z = 2819, 2820, 3073, 3332, 2567   # common_setting.f90:56-60
OBS_SD = {ID_T: 0.5, ID_S: 0.05, ID_U: 0.1, ID_V: 0.1, ID_Z: 0.03}
LON0, LAT0, DELTA = 130.0, 10.0, 1.0 / 16.0
RE = 6371.3e3


@dataclass
class Problem:
    params: dict
    lon1: object
    lat1: object
    ri1: object
    rj1: object
    dlon1: object
    dlat1: object
    phi1: object
    depth1: object      # (nij1, nlev)
    nobsgrd: object     # (nlon, nlat) int32
    obselm: object      # (nobs,) int32
    obslon: object
    obslat: object
    obslev: object
    obserr: object
    obsdep: object
    obshdxf: object     # (nobs, nmem)
    fcst3d: object      # (nij1, nlev, nmem, nv3d)
    fcst2d: object      # (nij1, nmem, nv2d)
    truth3d: object = None
    truth2d: object = None
    meta: dict = field(default_factory=dict)

    ARRAYS = ("lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1", "depth1",
              "nobsgrd", "obselm", "obslon", "obslat", "obslev", "obserr",
              "obsdep", "obshdxf", "fcst3d", "fcst2d")

    def numpy(self) -> "Problem":
        """Host copy with Fortran-ordered numpy arrays."""
        kw = dict(self.__dict__)
        for k in self.ARRAYS + ("truth3d", "truth2d"):
            v = kw[k]
            if isinstance(v, torch.Tensor):
                kw[k] = _to_fortran(v)
        return Problem(**kw)

    def to(self, device) -> "Problem":
        """Device copy: every array as a torch tensor whose memory is the
        Fortran layout (the tensor shape is the reversed Fortran shape)."""
        kw = dict(self.__dict__)
        for k in self.ARRAYS + ("truth3d", "truth2d"):
            v = kw[k]
            if isinstance(v, np.ndarray):
                kw[k] = torch.from_numpy(np.ascontiguousarray(v.T)).to(device)
            elif isinstance(v, torch.Tensor):
                kw[k] = v.to(device)
        return Problem(**kw)

    @property
    def dtype(self):
        v = self.fcst3d
        return v.dtype


def _to_fortran(t: torch.Tensor) -> np.ndarray:
    """A tensor whose C-order shape is the reversed Fortran shape -> the
    Fortran-ordered numpy array with the Fortran shape."""
    a = t.detach().cpu().numpy()
    return np.asfortranarray(a.T)


def _r(x: float, real_kind: int) -> float:
    return float(np.float32(x)) if real_kind == 4 else float(x)


def default_params(real_kind: int, sigma_obs: float, sigma_obsv: float,
                   alpha_rtpp: float = 0.9, alpha_rtps: float = 0.0,
                   cov_infl_mul: float = 1.0, roff: int = 1) -> dict:
    """The PARAMETERs of common_setting.f90:27-131 as r_size values."""
    R = lambda x: _r(x, real_kind)
    so, sv = R(sigma_obs), R(sigma_obsv)
    c = R(2.0 * math.sqrt(10.0 / 3.0))
    if real_kind == 4:
        dz = float(np.float32(c) * np.float32(so))
        dzv = float(np.float32(c) * np.float32(sv))
    else:
        dz, dzv = c * so, c * sv
    return dict(real_kind=real_kind, iv3d_t=3, iv3d_s=4, id_z_obs=ID_Z, roff=roff,
                pi=R(4.0 * math.atan(1.0)), re=R(RE), cov_infl_mul=R(cov_infl_mul),
                alpha_rtpp=R(alpha_rtpp), alpha_rtps=R(alpha_rtps),
                sigma_obs=so, sigma_obsv=sv, dist_zero=dz, dist_zerov=dzv,
                tmin=R(-1.8), tmax=R(50.0), smin=R(25.0), smax=R(50.0))


def _kernel(std: float, device) -> torch.Tensor:
    rad = max(1, int(math.ceil(3.0 * std)))
    x = torch.arange(-rad, rad + 1, dtype=torch.float64, device=device)
    w = torch.exp(-0.5 * (x / std) ** 2)
    return w / w.sum()


def _blur(x: torch.Tensor, dim: int, w: torch.Tensor) -> torch.Tensor:
    """Circular convolution of x along dim with the odd symmetric kernel w
    (shift-and-add: a handful of elementwise passes, no convolution library)."""
    rad = (w.numel() - 1) // 2
    n = x.shape[dim]
    wl = [float(v) for v in w.cpu()]
    out = x * wl[rad]
    for k in range(1, rad + 1):
        out += wl[rad + k] * (torch.roll(x, k % n, dims=dim) + torch.roll(x, -k % n, dims=dim))
    return out


def grf(shape, corr_h: float, corr_v: float, gen: torch.Generator, device) -> torch.Tensor:
    """Unit-variance Gaussian random field on (nlev, nlat, nlon) or
    (nlat, nlon); correlation exp(-r^2/L^2) with L = corr_h cells
    horizontally and corr_v levels vertically."""
    x = torch.randn(shape, generator=gen, dtype=torch.float32, device=device)
    var = 1.0
    wh = _kernel(corr_h / 2.0, device)
    x = _blur(x, -1, wh)
    x = _blur(x, -2, wh)
    var *= float((wh ** 2).sum()) ** 2
    if len(shape) == 3 and corr_v > 0 and shape[0] > 1:
        wv = _kernel(corr_v / 2.0, device)
        x = _blur(x, 0, wv)
        var *= float((wv ** 2).sum())
    return (x / math.sqrt(var)).double()


def make_problem(nlon=64, nlat=64, nlev=16, nmem=20, nv3d=4, nv2d=1, nobs=2000,
                 obs_layout="random", n_surface=0, n_profile_cols=0,
                 alpha_rtpp=0.7, alpha_rtps=0.0, cov_infl_mul=1.0, real_kind=8,
                 seed=20261203, corr_h=8.0, corr_v=2.0, land_frac=0.0,
                 sigma_h_cells=3.0, sigma_v=2.0, dz=1.0, device="cpu",
                 tracer=False, keep_truth=False, progress=None) -> Problem:
    """Synthetic problem.  obs_layout "random": nobs obs at uniform random
    nodes/levels with the element uniform over {U,V,T,S,SSH};
    "swath_profile": n_surface SST/SSH obs on 20-cell satellite-like stripes
    plus n_profile_cols full-depth T or S profiles at random columns."""
    dev = torch.device(device)
    rdt = torch.float32 if real_kind == 4 else torch.float64
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    nij1 = nlon * nlat
    # ---- grid (common_pom.f90 set_common_pom / set_common_mpi_pom) ----
    ii = np.arange(1, nlon + 1, dtype=np.float64)
    jj = np.arange(1, nlat + 1, dtype=np.float64)
    lon = LON0 + (ii - 1) * DELTA
    lat = LAT0 + (jj - 1) * DELTA
    dy = RE * DELTA * math.pi / 180.0
    dx = dy * np.cos(lat * math.pi / 180.0)
    LON, LAT = np.meshgrid(lon, lat)          # (nlat, nlon)
    DX = np.broadcast_to(dx[:, None], (nlat, nlon))
    IX, JX = np.meshgrid(ii, jj)
    phi = np.full((nlat, nlon), 5000.0)
    if land_frac > 0:
        phi[rng.random((nlat, nlon)) < land_frac] = 1.0
    depth_k = -(nlev - np.arange(1, nlev + 1, dtype=np.float64)) * dz   # k = nlev surface
    params = default_params(real_kind, sigma_h_cells * dy, sigma_v, alpha_rtpp,
                            alpha_rtps, cov_infl_mul)

    def R(a):
        return torch.as_tensor(np.ascontiguousarray(a), device=dev).to(rdt)

    lon1 = R(LON.ravel())
    lat1 = R(LAT.ravel())
    ri1 = R(IX.ravel())
    rj1 = R(JX.ravel())
    dlon1 = R(DX.ravel())
    dlat1 = R(np.full(nij1, dy))
    phi1 = R(phi.ravel())
    depth1 = R(np.broadcast_to(depth_k[:, None], (nlev, nij1)))    # memory (nij1,nlev) F
    # ---- truth and ensemble ----
    names3 = ["U", "V", "T", "S"] + (["TRACER"] if tracer else [])
    names3 = names3[:nv3d] if nv3d <= len(names3) else names3 + ["X%d" % q for q in range(nv3d - len(names3))]
    mean3 = {"U": 0.0, "V": 0.0, "T": 15.0, "S": 34.5}
    sd3 = {"U": 0.2, "V": 0.2, "T": 3.0, "S": 0.3}
    truth3 = torch.empty((nv3d, nlev, nlat, nlon), dtype=torch.float64, device=dev)
    g3 = []
    for v, nm in enumerate(names3):
        g = grf((nlev, nlat, nlon), corr_h, corr_v, gen, dev)
        g3.append(g)
        truth3[v] = torch.exp(0.5 * g) if nm == "TRACER" else mean3.get(nm, 0.0) + sd3.get(nm, 1.0) * g
    truth2 = torch.empty((nv2d, nlat, nlon), dtype=torch.float64, device=dev)
    g2 = []
    for v in range(nv2d):
        g = grf((nlat, nlon), corr_h, 0.0, gen, dev)
        g2.append(g)
        truth2[v] = 0.1 * g
    fcst3 = torch.empty((nv3d, nmem, nlev, nlat, nlon), dtype=rdt, device=dev)
    fcst2 = torch.empty((nv2d, nmem, nlat, nlon), dtype=rdt, device=dev)
    for v, nm in enumerate(names3):
        if progress:
            progress(f"ensemble variable {nm} ({v + 1}/{nv3d})")
        for m in range(nmem):
            e = grf((nlev, nlat, nlon), corr_h, corr_v, gen, dev)
            if nm == "TRACER":
                fcst3[v, m] = torch.exp(0.5 * (g3[v] + 0.5 * e)).to(rdt)
            else:
                fcst3[v, m] = (truth3[v] + 0.5 * sd3.get(nm, 1.0) * e).to(rdt)
    for v in range(nv2d):
        for m in range(nmem):
            e = grf((nlat, nlon), corr_h, 0.0, gen, dev)
            fcst2[v, m] = (truth2[v] + 0.05 * e).to(rdt)
    # ---- observations on grid nodes ----
    vidx = {ID_U: 0, ID_V: 1, ID_T: 2, ID_S: 3}
    if obs_layout == "random":
        oi = rng.integers(1, nlon, nobs)          # 1..nlon-1 (cell exists)
        oj = rng.integers(1, nlat, nobs)
        ok = rng.integers(1, nlev + 1, nobs)
        choices = [e for e in (ID_U, ID_V, ID_T, ID_S) if vidx[e] < nv3d]
        if nv2d > 0:
            choices.append(ID_Z)
        oe = np.asarray(choices, np.int32)[rng.integers(0, len(choices), nobs)]
        ok = np.where(oe == ID_Z, nlev, ok)
    elif obs_layout == "swath_profile":
        # stripes 20 cells wide, slanted like satellite swaths
        ci, cj = np.meshgrid(np.arange(1, nlon), np.arange(1, nlat))
        in_stripe = ((ci - cj // 2) % 64) < 20
        cells = np.flatnonzero(in_stripe.ravel())
        ncell = min(len(cells), n_surface // 2)
        pick = rng.choice(cells, ncell, replace=False)
        si, sj = ci.ravel()[pick], cj.ravel()[pick]
        surf_i = np.repeat(si, 2)
        surf_j = np.repeat(sj, 2)
        surf_e = np.tile(np.array([ID_T, ID_Z], np.int32), ncell)
        if nv2d == 0:
            surf_e[:] = ID_T
        surf_k = np.full(2 * ncell, nlev)
        cols = rng.choice((nlon - 1) * (nlat - 1), n_profile_cols, replace=False)
        pi_, pj_ = cols % (nlon - 1) + 1, cols // (nlon - 1) + 1
        pe = np.where(np.arange(n_profile_cols) % 2 == 0, ID_T, ID_S).astype(np.int32)
        prof_i = np.repeat(pi_, nlev)
        prof_j = np.repeat(pj_, nlev)
        prof_k = np.tile(np.arange(1, nlev + 1), n_profile_cols)
        prof_e = np.repeat(pe, nlev)
        oi = np.concatenate([surf_i, prof_i])
        oj = np.concatenate([surf_j, prof_j])
        ok = np.concatenate([surf_k, prof_k])
        oe = np.concatenate([surf_e, prof_e]).astype(np.int32)
        order = rng.permutation(len(oi))          # input (file) order is arbitrary
        oi, oj, ok, oe = oi[order], oj[order], ok[order], oe[order]
    else:
        raise ValueError(obs_layout)
    n_o = len(oi)
    # H(x_m): point sampling at the node
    ti = torch.as_tensor(oi - 1, device=dev)
    tj = torch.as_tensor(oj - 1, device=dev)
    tk = torch.as_tensor(ok - 1, device=dev)
    te = torch.as_tensor(oe, device=dev)
    hx = torch.empty((nmem, n_o), dtype=torch.float64, device=dev)
    tru = torch.empty(n_o, dtype=torch.float64, device=dev)
    for e, v in list(vidx.items()) + [(ID_Z, -1)]:
        sel = torch.nonzero(te == e).flatten()
        if sel.numel() == 0:
            continue
        if e == ID_Z:
            hx[:, sel] = fcst2[0][:, tj[sel], ti[sel]].double()
            tru[sel] = truth2[0][tj[sel], ti[sel]]
        else:
            hx[:, sel] = fcst3[v][:, tk[sel], tj[sel], ti[sel]].double()
            tru[sel] = truth3[v][tk[sel], tj[sel], ti[sel]]
    sd = torch.as_tensor(np.array([OBS_SD[int(e)] for e in oe]), device=dev)
    noise = torch.randn(n_o, generator=gen, dtype=torch.float64, device=dev)
    yo = tru + sd * noise
    # set_obs (common_obs.f90:144-181): mean, dYf, innovation, in r_size
    hxr = hx.to(rdt)
    hmean = hxr.sum(0) / nmem
    hdxf = (hxr - hmean[None, :])
    dep = (yo.to(rdt) - hmean)
    err = sd.to(rdt)
    olon = torch.as_tensor(LON0 + (oi - 1) * DELTA, device=dev).to(rdt)
    olat = torch.as_tensor(LAT0 + (oj - 1) * DELTA, device=dev).to(rdt)
    olev = torch.as_tensor(depth_k[ok - 1], device=dev).to(rdt)
    # bucket sort by (lat cell, lon cell), stable (common_obs.f90:266-377)
    key = (oj - 1) * nlon + (oi - 1)
    perm = np.argsort(key, kind="stable")
    counts = np.bincount(key, minlength=nlon * nlat)
    cum = np.cumsum(counts).astype(np.int64)
    grd = cum.reshape(nlat, nlon)
    grd[nlat - 1, :] = 0                      # row nlat is never filled
    tp = torch.as_tensor(perm, device=dev)
    prob = Problem(
        params=dict(params),
        lon1=lon1, lat1=lat1, ri1=ri1, rj1=rj1, dlon1=dlon1, dlat1=dlat1, phi1=phi1,
        depth1=depth1,
        nobsgrd=torch.as_tensor(grd.astype(np.int32), device=dev),
        obselm=te[tp].to(torch.int32).contiguous(),
        obslon=olon[tp].contiguous(), obslat=olat[tp].contiguous(),
        obslev=olev[tp].contiguous(), obserr=err[tp].contiguous(),
        obsdep=dep[tp].contiguous(),
        obshdxf=hdxf[:, tp].contiguous(),          # memory (nobs,nmem) F
        fcst3d=fcst3.reshape(nv3d, nmem, nlev, nij1).contiguous(),
        fcst2d=fcst2.reshape(nv2d, nmem, nij1).contiguous(),
        meta=dict(nlon=nlon, nlat=nlat, nlev=nlev, nmem=nmem, nv3d=nv3d, nv2d=nv2d,
                  nobs=n_o, seed=seed, obs_layout=obs_layout))
    if keep_truth:
        prob.truth3d = truth3
        prob.truth2d = truth2
    return prob


# Workload c4: eight independent analysis cycles, grids 128^2 .. 512^2 x 40,
# nmem round-robin over (32, 64, 100, 128), 0.5 .. 4 random obs per column.
C4_GRIDS = (128, 192, 256, 256, 320, 384, 448, 512)
C4_NMEM = (32, 64, 100, 128)
C4_DENSITY = (0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0)


def c4_cycles(scale=1.0, nlev=40):
    """make_problem arguments of the eight c4 cycles; `scale` shrinks the
    horizontal grid (the parity test runs them at a few % of full size)."""
    out = []
    for c, (n, d) in enumerate(zip(C4_GRIDS, C4_DENSITY)):
        m = max(8, int(round(n * scale)))
        out.append(dict(nlon=m, nlat=m, nlev=nlev, nmem=C4_NMEM[c % 4], nv3d=4, nv2d=1,
                        nobs=max(1, int(d * m * m)), obs_layout="random", alpha_rtpp=0.9,
                        seed=20261207 + c))
    return out


def config(name: str, device="cpu", real_kind=8, **kw) -> Problem:
    """Named workloads: c0 (64x64x16, K=20, 2000 random obs, RTPP 0.7),
    c1 (512x512x40, K=100, 6 vars incl. tracer, 1.5e5 swath + 5e4 profile
    obs), c2 (c1 obs geometry, K=64, U/V/T/S + SSH), c3 (256x256x32, K=32,
    4 vars, 5e4 random obs)."""
    if name == "c0":
        a = dict(nlon=64, nlat=64, nlev=16, nmem=20, nv3d=4, nv2d=1, nobs=2000,
                 alpha_rtpp=0.7)
    elif name == "c1":
        a = dict(nlon=512, nlat=512, nlev=40, nmem=100, nv3d=5, nv2d=1, tracer=True,
                 obs_layout="swath_profile", n_surface=150000, n_profile_cols=1250,
                 alpha_rtpp=0.9)
    elif name == "c2":
        a = dict(nlon=512, nlat=512, nlev=40, nmem=64, nv3d=4, nv2d=1,
                 obs_layout="swath_profile", n_surface=150000, n_profile_cols=1250,
                 alpha_rtpp=0.9)
    elif name == "c3":
        a = dict(nlon=256, nlat=256, nlev=32, nmem=32, nv3d=4, nv2d=0, nobs=50000,
                 alpha_rtpp=0.9, corr_h=6.0)
    else:
        raise ValueError(name)
    a.update(kw)
    return make_problem(device=device, real_kind=real_kind, **a)
