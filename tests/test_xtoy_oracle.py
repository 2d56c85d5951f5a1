"""The observation operator H(x) of the oracle (oracle_trans_xtoy /
oracle_akima, restating Trans_XtoY and akima_spline of
letkf/run/src/common.f90:246-500) pinned against the reference's own
compiled routines (oracle/_ref, FP32 as shipped, its fixed 1442 x 562 x 75
grid) and against closed forms."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
import xtoy_cases as xc

REF = po.ref()
needs_ref = pytest.mark.skipif(REF is None, reason="reference build (oracle/_ref) not present")


def _xg(dims, hnosea=0.0, undef=999.0):
    return po.OracleXtoY(*dims, hnosea, undef)


def _F(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@needs_ref
def test_akima_matches_reference_fp32():
    rng = np.random.default_rng(5)
    nbad = 0
    for trial in range(400):
        ndim = int(rng.integers(5, 40))
        x = np.cumsum(rng.uniform(0.1, 5.0, ndim)).astype(np.float32) - 200.0
        y = (np.cumsum(rng.standard_normal(ndim)) * rng.uniform(0.1, 3)).astype(np.float32)
        if trial % 5 == 0:
            y = (0.3 * x + 2.0).astype(np.float32)      # linear: both Akima weights vanish
        targets = list(rng.uniform(x[0], x[-1], 6).astype(np.float32)) + [x[0], x[int(rng.integers(1, ndim - 1))],
                                                                         x[0] - 1.0, x[-1] + 1.0]
        for t in targets:
            out = C.c_float()
            REF.ref_akima(ndim, _F(x), _F(y), C.c_float(float(t)), C.byref(out))
            mine = po.akima(x, y, float(t), dtype=np.float32)
            if np.float32(out.value) != np.float32(mine):
                nbad += 1
    assert nbad == 0


@needs_ref
def test_trans_xtoy_matches_reference_fp32():
    dims = (C.c_int32 * 15)()
    vals = (C.c_float * 2)()
    REF.ref_xtoy_consts(dims, vals)
    nlon, nlat, nlev, nv3d, nv2d = dims[:5]
    g = _xg(list(dims), float(vals[0]), float(vals[1]))
    rng = np.random.default_rng(11)
    region = ((300, 360), (200, 250))
    lon, lat, fsm, depth = xc.make_grid(rng, nlon, nlat, nlev, region, np.float32)
    v3d, v2d = xc.make_fields(rng, 1, nv3d, nv2d, nlev, nlat, nlon, region, np.float32)
    v3d, v2d = v3d[0], v2d[0]
    obs = xc.make_obs(rng, 3000, lon, lat, depth, region, nlon, nlat, np.float32)
    elm, idx, idy, lono, lato, levo = obs
    hx_ref = np.empty(len(elm), np.float32)
    IP = C.POINTER(C.c_int32)
    REF.ref_trans_xtoy(len(elm), elm.ctypes.data_as(IP), idx.ctypes.data_as(IP), idy.ctypes.data_as(IP),
                       _F(lono), _F(lato), _F(levo), _F(v3d), _F(v2d), _F(lon), _F(lat), _F(depth), _F(fsm),
                       _F(hx_ref))
    # the oracle takes Fortran-ordered arrays: numpy's reversed axes already are
    hx = po.trans_xtoy(g, elm, idx, idy, lono, lato, levo, lon.T, lat.T, depth.T, fsm.T, v3d.T, v2d.T,
                       dtype=np.float32)
    undef = np.float32(999.0)
    # Outside the reference's defined behaviour, and excluded from the
    # comparison: a target equal to a column's top level (its Akima search
    # reads past the column; it returned NaN here), and an element
    # Trans_XtoY does not know (it reuses iv3d from the obs processed before,
    # uninitialised per OpenMP thread).  For those the oracle returns the top
    # level's value and undef.
    top_hit = np.zeros(len(elm), bool)
    ii, jj = np.clip(idx, 1, nlon - 1) - 1, np.clip(idy, 1, nlat - 1) - 1
    for di, dj in ((0, 0), (1, 0), (0, 1), (1, 1)):
        top_hit |= levo == depth[nlev - 1, jj + dj, ii + di]
    unknown = elm == 9999
    keep = ~top_hit & ~unknown
    assert (hx_ref == undef).sum() > 100 and (hx_ref != undef).sum() > 1000 and keep.sum() > 2800
    assert np.array_equal(hx[keep], hx_ref[keep])
    assert np.all(hx[unknown] == undef)


def test_akima_closed_forms():
    # linear data: dydx equal, both weights zero -> t = 0 in the reference,
    # so y5 = y(i-1) + a*dx21*(3 th^2 - 2 th^3), th = dx/dx21
    x = np.array([-50.0, -40.0, -30.0, -20.0, -10.0, 0.0])
    y = 2.0 * x + 5.0
    for x5 in (-47.0, -33.0, -21.5, -2.0):
        i = int(np.searchsorted(x, x5, side="right"))
        dx21, th = x[i] - x[i - 1], (x5 - x[i - 1]) / (x[i] - x[i - 1])
        want = y[i - 1] + 2.0 * dx21 * (3 * th ** 2 - 2 * th ** 3)
        assert abs(po.akima(x, y, x5) - want) < 1e-12
    # exact hits (including the top level) return the data; outside -> undef
    for k in range(6):
        assert po.akima(x, y, x[k]) == y[k]
    assert po.akima(x, y, -60.0) == 999.0 and po.akima(x, y, 1.0) == 999.0


def test_trans_xtoy_closed_forms_fp64():
    rng = np.random.default_rng(3)
    nlon, nlat, nlev, nv3d, nv2d = 20, 16, 12, 4, 2
    region = ((1, nlon), (1, nlat))
    lon, lat, fsm, depth = xc.make_grid(rng, nlon, nlat, nlev, region)
    fsm[:] = 1.0
    g = _xg([nlon, nlat, nlev, nv3d, nv2d, xc.ID_U, xc.ID_V, xc.ID_T, xc.ID_S, xc.ID_Z, 1, 2, 3, 4, 1])
    # SSH field bilinear in (lon, lat): H reproduces it
    v2d = np.zeros((nv2d, nlat, nlon))
    v2d[0] = 1.5 + 0.2 * lon - 0.3 * lat + 0.01 * lon * lat
    v3d = np.zeros((nv3d, nlev, nlat, nlon))
    v3d[2] = 7.0  # constant temperature: every branch that interpolates returns it
    n = 200
    idx = rng.integers(1, nlon - 1, n).astype(np.int32)
    idy = rng.integers(1, nlat - 1, n).astype(np.int32)
    lono = lon[idy - 1, idx - 1] + rng.uniform(0, 0.25, n)
    lato = lat[idy - 1, idx - 1] + rng.uniform(0, 0.25, n)
    elm = np.full(n, xc.ID_Z, np.int32)
    hx = po.trans_xtoy(g, elm, idx, idy, lono, lato, np.zeros(n), lon.T, lat.T, depth.T, fsm.T, v3d.T, v2d.T)
    want = 1.5 + 0.2 * lono - 0.3 * lato + 0.01 * lono * lato
    assert np.abs(hx - want).max() < 1e-11
    # constant field, interior depths: Akima (t = 0 on constant data) and bilinear give it back
    elm[:] = xc.ID_T
    ii, jj = idx - 1, idy - 1
    top = np.min([depth[nlev - 1, jj + a, ii + b] for a in (0, 1) for b in (0, 1)], axis=0)
    bot = np.max([depth[1, jj + a, ii + b] for a in (0, 1) for b in (0, 1)], axis=0)
    levo = bot + rng.uniform(0.05, 0.95, n) * (top - bot)
    hx = po.trans_xtoy(g, elm, idx, idy, lono, lato, levo, lon.T, lat.T, depth.T, fsm.T, v3d.T, v2d.T)
    assert np.abs(hx - 7.0).max() < 1e-12
    # land corner / above surface / outside the grid -> undef
    fsm2 = fsm.copy()
    fsm2[jj[0], ii[0] + 1] = 0.0
    hx = po.trans_xtoy(g, elm[:1], idx[:1], idy[:1], lono[:1], lato[:1], levo[:1], lon.T, lat.T, depth.T,
                       fsm2.T, v3d.T, v2d.T)
    assert hx[0] == 999.0
    hx = po.trans_xtoy(g, elm[:1], idx[:1], idy[:1], lono[:1], lato[:1], np.array([0.5]), lon.T, lat.T,
                       depth.T, fsm.T, v3d.T, v2d.T)
    assert hx[0] == 999.0
    hx = po.trans_xtoy(g, elm[:1], np.array([nlon], np.int32), idy[:1], lono[:1], lato[:1], levo[:1], lon.T,
                       lat.T, depth.T, fsm.T, v3d.T, v2d.T)
    assert hx[0] == 999.0
