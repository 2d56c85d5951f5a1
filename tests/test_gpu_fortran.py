"""The Fortran side of the boundary, executed: oracle/_ref/letkf_dropin (built
from tests/fortran/dropin_driver.f90, the engine's iso_c_binding module
sbpom_letkf_v2_amd/fortran/letkf_engine_mod.f90 and the reference's own
common_setting.f90, unmodified) fills the reference's module state from a
file and calls letkf_da_main_gpu twice, as main.f90:92 calls da_main.

The problem has the reference's compiled-in sizes (REAL(4), nmem 128,
nlon x nlat 1442 x 562, nlev 75, nv3d 4, nv2d 3) on a 6 x 5 column patch
of the global grid, with its localisation (300 km, 100 m), RTPP 0.9 and
round_off.  Both analyses (FP64 Gram plain form, then FP32 MFMA Gram
batched form through letkf_set_params on the same handle) must match the
FP32-build oracle within 2e-5 of the field scale, the REAL(4) bar.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from helpers import per_var_err
from sbpom_letkf_v2_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "letkf_dropin")

NLON, NLAT, NLEV, K, NV3, NV2 = 1442, 562, 75, 128, 4, 3   # common_setting.f90:36-42


def _problem(seed=5, nobs=400, i0=700, j0=280, pw=6, ph=5):
    rng = np.random.default_rng(seed)
    f32 = np.float32
    d = synth.DELTA
    ii, jj = np.meshgrid(np.arange(i0, i0 + pw), np.arange(j0, j0 + ph))   # 1-based global cells
    ii, jj = ii.ravel(), jj.ravel()
    nij1 = ii.size
    lat = synth.LAT0 + (jj - 1) * d
    dy = synth.RE * d * np.pi / 180.0
    grid = dict(lon1=synth.LON0 + (ii - 1) * d, lat1=lat, ri1=ii.astype(float), rj1=jj.astype(float),
                dlon1=dy * np.cos(lat * np.pi / 180.0), dlat1=np.full(nij1, dy),
                phi1=np.where(rng.random(nij1) < 0.15, 1.0, 3000.0))
    grid = {k: v.astype(f32) for k, v in grid.items()}
    depth_k = (-(NLEV - np.arange(1, NLEV + 1)) * 20.0).astype(f32)      # k = nlev the surface
    grid["depth1"] = np.asfortranarray(np.broadcast_to(depth_k[None, :], (nij1, NLEV)).astype(f32))
    # obs on nodes around the patch, sorted by (lat cell, lon cell)
    oi = rng.integers(i0 - 12, i0 + pw + 12, nobs)
    oj = rng.integers(j0 - 12, j0 + ph + 12, nobs)
    elems = np.array([synth.ID_U, synth.ID_V, synth.ID_T, synth.ID_S, synth.ID_Z], np.int32)
    oe = elems[rng.integers(0, 5, nobs)]
    ok = np.where(oe == synth.ID_Z, NLEV, rng.integers(1, NLEV + 1, nobs))
    order = np.argsort((oj - 1) * NLON + (oi - 1), kind="stable")
    oi, oj, oe, ok = oi[order], oj[order], oe[order], ok[order]
    counts = np.bincount((oj - 1) * NLON + (oi - 1), minlength=NLON * NLAT)
    grd = np.cumsum(counts).reshape(NLAT, NLON)
    grd[NLAT - 1, :] = 0
    sd = np.array([synth.OBS_SD[int(e)] for e in oe])
    hdxf = rng.standard_normal((nobs, K)) * sd[:, None] * 2.0
    hdxf -= hdxf.mean(1, keepdims=True)
    obs = dict(nobsgrd=np.asfortranarray(grd.T.astype(np.int32)), obselm=oe,
               obslon=(synth.LON0 + (oi - 1) * d).astype(f32), obslat=(synth.LAT0 + (oj - 1) * d).astype(f32),
               obslev=depth_k[ok - 1], obserr=sd.astype(f32), obsdep=(rng.standard_normal(nobs) * sd).astype(f32),
               obshdxf=np.asfortranarray(hdxf.astype(f32)))
    mean3 = np.array([0.0, 0.0, 15.0, 34.5])
    sd3 = np.array([0.2, 0.2, 3.0, 0.3])
    x = rng.standard_normal((nij1, NLEV, 1, NV3)) + 0.5 * rng.standard_normal((nij1, NLEV, K, NV3))
    f3 = np.asfortranarray((mean3 + sd3 * x).astype(f32))
    f2 = np.asfortranarray((0.1 * rng.standard_normal((nij1, 1, NV2)) +
                            0.05 * rng.standard_normal((nij1, K, NV2))).astype(f32))
    params = synth.default_params(4, 300e3, 100.0, alpha_rtpp=0.9)
    return po.Problem(params=params, fcst3d=f3, fcst2d=f2, **grid, **obs)


def _write(path, pr):
    F = lambda a: np.asarray(a).ravel(order="F")
    with open(path, "wb") as f:
        np.array([pr.lon1.size, pr.obselm.size], np.int32).tofile(f)
        for a in (pr.lon1, pr.lat1, pr.ri1, pr.rj1, pr.dlon1, pr.dlat1, pr.phi1, pr.depth1):
            F(a).astype(np.float32).tofile(f)
        F(pr.nobsgrd).astype(np.int32).tofile(f)
        F(pr.obselm).astype(np.int32).tofile(f)
        for a in (pr.obslon, pr.obslat, pr.obslev, pr.obserr, pr.obsdep, pr.obshdxf, pr.fcst3d, pr.fcst2d):
            F(a).astype(np.float32).tofile(f)


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="oracle/_ref/letkf_dropin not built (needs amdflang and the reference's "
                           "common_setting.f90 at build time)")
def test_fortran_dropin_runs_da_main_on_the_gpu(tmp_path):
    pr = _problem()
    nij1 = pr.lon1.size
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write(fin, pr)
    print(f"case Fortran drop-in: nij1={nij1} nlev={NLEV} nmem={K} nobs={pr.obselm.size}", flush=True)
    r = subprocess.run([DROPIN, str(fin), str(fout)], capture_output=True, text=True, timeout=240)
    print(r.stdout + r.stderr, flush=True)
    assert r.returncode == 0
    out = np.fromfile(fout, np.float32)
    n3, n2 = nij1 * NLEV * K * NV3, nij1 * K * NV2
    assert out.size == 2 * (n3 + n2)
    rc, a3o, a2o, nl = po.da_main(pr, dtype=np.float32)
    assert rc == 0
    print(f"  nobsl {nl.min()}..{nl.max()}", flush=True)
    assert nl.max() >= K and nl.min() < K      # both the K x K and the n x n forms
    o = 0
    for name in ("fp64/plain", "fp32/batched"):
        a3 = out[o:o + n3].reshape((nij1, NLEV, K, NV3), order="F")
        o += n3
        a2 = out[o:o + n2].reshape((nij1, K, NV2), order="F")
        o += n2
        e3 = per_var_err(a3, a3o, axis_var=3)
        e2 = per_var_err(a2, a2o, axis_var=2)
        print(f"  {name}: err3={max(e3):.2e} err2={max(e2):.2e}", flush=True)
        assert max(e3) <= 2e-5 and max(e2) <= 2e-5
