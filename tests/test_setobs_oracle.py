"""Pins the oracle's set_obs (letkf/run/src/common_obs.f90:139-402).

The reference's set_obs needs MPI and NetCDF, so it is not built here; the
oracle is checked step by step against what the algorithm fixes:
  * compaction + (lat cell, lon cell) sort == numpy's stable argsort of the
    cell key over the kept obs, nobsgrd == cumulative bincount (row nlat 0);
  * mean / spread against math.fsum, hdxf and dep as exact differences,
    the one-member and identical-member special cases;
  * gross-error decisions at the reference's bounds (common_setting.f90:
    109-118) on hand-placed innovations, and the AOEI switch that disables
    them (common_obs.f90:185);
  * temporal localisation factors exp(0.25 ((islot-nbslot)/sigma_obst)^2);
  * AOEI as sqrt(max(err^2, dep^2 - sprd^2)) through its two branches.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from setobs_cases import ID_S, ID_T, ID_U, ID_V, ID_Z, UNKNOWN, make_case


def run(c, case, dtype):
    return po.set_obs(c, case["elm"], case["lon"], case["lat"], case["lev"], case["dat"], case["err"],
                      case["idx"], case["idy"], case["hx"], qc=case["qc"], ins=case["ins"],
                      nobslots=case["nobslots"], dtype=dtype)


def source_rows(case, res):
    """Input index of every sorted row, from the unique longitudes."""
    where = {float(v): i for i, v in enumerate(case["lon"])}
    return np.array([where[float(v)] for v in res["lon"]], dtype=np.int64)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sort_and_nobsgrd_match_stable_argsort(dtype):
    rng = np.random.default_rng(11)
    nlon, nlat, nmem = 13, 9, 6
    case = make_case(rng, 900, nmem, nlon, nlat, dtype)
    c = po.setobs_config(nmem, nlon, nlat)
    res = run(c, case, dtype)
    ix, iy = case["idx"], case["idy"]
    inside = (ix >= 1) & (ix <= nlon - 1) & (iy >= 1) & (iy <= nlat - 1)
    keep = np.flatnonzero((case["qc"] == 1) & inside)
    key = (iy[keep] - 1) * nlon + (ix[keep] - 1)
    want = keep[np.argsort(key, kind="stable")]
    got = source_rows(case, res)
    assert res["nobs"] == len(keep)
    np.testing.assert_array_equal(got, want)
    cnt = np.bincount(key, minlength=nlon * nlat)
    grd = np.cumsum(cnt).reshape(nlat, nlon).T.copy()
    grd[:, nlat - 1] = 0
    np.testing.assert_array_equal(res["nobsgrd"], grd)
    assert res["ndel"] == (900 - len(keep), 0)
    for k in ("elm", "idx", "idy", "ins", "lat", "lev", "dat"):
        np.testing.assert_array_equal(res[k], case[k][got])
    # sorted by cell: (idy, idx) non-decreasing
    kk = (res["idy"].astype(np.int64) - 1) * nlon + res["idx"] - 1
    assert np.all(np.diff(kk) >= 0)


def test_statistics_closed_forms():
    rng = np.random.default_rng(12)
    nlon, nlat, nmem = 10, 8, 7
    case = make_case(rng, 400, nmem, nlon, nlat)
    c = po.setobs_config(nmem, nlon, nlat, aoei=0)
    res = run(c, case, np.float64)
    src = source_rows(case, res)
    hx = case["hx"][src]
    mean = np.array([math.fsum(r) / nmem for r in hx])
    np.testing.assert_allclose(res["hxfmean"], mean, rtol=1e-15, atol=0)
    sprd = np.array([math.sqrt(math.fsum((r - m) ** 2) / (nmem - 1)) for r, m in zip(hx, res["hxfmean"])])
    np.testing.assert_allclose(res["hxfsprd"], sprd, rtol=1e-13, atol=0)
    # perturbations and innovations are single differences from the stored mean
    np.testing.assert_array_equal(res["hdxf"], hx - res["hxfmean"][:, None])
    np.testing.assert_array_equal(res["dep"], case["dat"][src] - res["hxfmean"])
    np.testing.assert_array_equal(res["err"], case["err"][src])  # one slot, no AOEI


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_identical_members_and_single_member(dtype):
    rng = np.random.default_rng(13)
    nlon, nlat = 9, 7
    case = make_case(rng, 120, 5, nlon, nlat, dtype, p_qc0=0.0, p_out=0.0)
    # identical members on a 1/64 lattice: the member sum and the division
    # by nmem are exact, so the mean is the member value itself
    case["hx"] = np.repeat(np.round(case["hx"][:, :1] * 64) / 64, 5, axis=1).astype(dtype)
    res = run(po.setobs_config(5, nlon, nlat, aoei=0), case, dtype)
    assert res["nobs"] == 120
    assert np.all(res["hxfsprd"] == 0) and np.all(res["hdxf"] == 0)
    src = source_rows(case, res)
    np.testing.assert_array_equal(res["hxfmean"], case["hx"][src, 0])
    # nmem == 1: spread is undef (common_obs.f90:159-163)
    one = dict(case, hx=case["hx"][:, :1].copy())
    res1 = run(po.setobs_config(1, nlon, nlat, aoei=1), one, dtype)
    assert np.all(res1["hxfsprd"] == 999.0) and np.all(res1["hdxf"] == 0)
    # AOEI with sprd = undef: dep^2 - 999^2 < err^2 for these innovations
    np.testing.assert_array_equal(res1["err"], case["err"][src])


def _gross_case(rows, dtype):
    """rows of (elm, lev, dep, err); hx constant = dat - dep."""
    n = len(rows)
    elm = np.array([r[0] for r in rows], np.int32)
    lev = np.array([r[1] for r in rows], dtype)
    dep = np.array([r[2] for r in rows], np.float64)
    err = np.array([r[3] for r in rows], dtype)
    dat = np.full(n, 20.0)
    hx = np.repeat((dat - dep)[:, None], 3, axis=1).astype(dtype)
    return dict(elm=elm, lon=np.arange(n, dtype=dtype) + 100, lat=np.zeros(n, dtype), lev=lev,
                dat=dat.astype(dtype), err=err, idx=np.ones(n, np.int32), idy=np.ones(n, np.int32),
                hx=hx, qc=np.ones(n, np.int32), ins=np.zeros(n, np.int32), nobslots=None)


GROSS_ROWS = [  # (elm, lev, innovation, err, kept?)  bounds: common_setting.f90:109-118
    (ID_Z, 0.0, 0.75, 0.1, True), (ID_Z, 0.0, 1.25, 0.1, False), (ID_Z, 0.0, -1.25, 0.1, False),
    (ID_T, 0.0, 4.5, 0.1, True), (ID_T, 0.0, -5.5, 0.1, False),       # SST
    (ID_T, -30.0, 4.75, 0.1, True), (ID_T, -30.0, 5.25, 0.1, False),  # T
    (ID_S, 0.0, 1.75, 0.1, True), (ID_S, 0.0, 2.25, 0.1, False),      # SSS
    (ID_S, -80.0, -2.5, 0.1, False), (ID_S, -80.0, -1.5, 0.1, True),  # S
    (ID_U, 0.0, 2.5, 0.1, False), (ID_U, -10.0, 1.5, 0.1, True),      # SSU, U
    (ID_V, 0.0, -1.75, 0.1, True), (ID_V, -10.0, 3.0, 0.1, False),    # SSV, V
    (UNKNOWN, -5.0, 9.5, 1.0, True), (UNKNOWN, -5.0, 10.5, 1.0, False),  # |dep| > 10 err
    (UNKNOWN, -5.0, -10.5, 1.0, False),
]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_gross_error_known_answers(dtype):
    case = _gross_case([r[:4] for r in GROSS_ROWS], dtype)
    n = len(GROSS_ROWS)
    kept = [i for i, r in enumerate(GROSS_ROWS) if r[4]]
    res = run(po.setobs_config(3, 4, 3, lge_io=1, aoei=0), case, dtype)
    assert res["nobs"] == len(kept)
    assert res["ndel"] == (0, n - len(kept))
    np.testing.assert_array_equal(source_rows(case, res), kept)
    # with AOEI on the gross check is skipped (common_obs.f90:185)
    res = run(po.setobs_config(3, 4, 3, lge_io=1, aoei=1), case, dtype)
    assert res["nobs"] == n and res["ndel"] == (0, 0)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_temporal_localisation_factors(dtype):
    rng = np.random.default_rng(14)
    nlon, nlat, nmem = 8, 6, 4
    case = make_case(rng, 300, nmem, nlon, nlat, dtype, nslots=4)
    c = po.setobs_config(nmem, nlon, nlat, aoei=0, nbslot=2, sigma_obst=2.0)
    res = run(c, case, dtype)
    src = source_rows(case, res)
    slot = np.searchsorted(np.cumsum(case["nobslots"]), src, side="right") + 1
    fac = np.array([math.exp(0.25 * ((s - 2) / 2.0) ** 2) for s in slot])
    want = (case["err"][src].astype(np.float64) * fac).astype(dtype)
    np.testing.assert_array_equal(res["err"], want)
    with pytest.raises(ValueError):
        run(c, dict(case, nobslots=case["nobslots"] + 1), dtype)


def test_aoei_branches():
    rng = np.random.default_rng(15)
    nlon, nlat, nmem = 11, 7, 8
    case = make_case(rng, 600, nmem, nlon, nlat)
    plain = run(po.setobs_config(nmem, nlon, nlat, aoei=0), case, np.float64)
    res = run(po.setobs_config(nmem, nlon, nlat, aoei=1), case, np.float64)
    for k in ("lon", "dep", "hdxf", "hxfsprd"):
        np.testing.assert_array_equal(res[k], plain[k])
    e0, d, s = plain["err"], plain["dep"], plain["hxfsprd"]
    grow = d * d - s * s > e0 * e0
    assert 0 < grow.sum() < len(grow)
    np.testing.assert_array_equal(res["err"][~grow], e0[~grow])
    np.testing.assert_allclose(res["err"][grow] ** 2, (d * d - s * s)[grow], rtol=1e-14)
    assert np.all(res["err"] >= e0)


def test_empty_and_all_rejected():
    rng = np.random.default_rng(16)
    c = po.setobs_config(3, 5, 4)
    case = make_case(rng, 50, 3, 5, 4)
    res = run(c, dict(case, qc=np.zeros(50, np.int32)), np.float64)
    assert res["nobs"] == 0 and res["ndel"] == (50, 0) and not res["nobsgrd"].any()
    empty = {k: (v[:0] if isinstance(v, np.ndarray) else v) for k, v in case.items()}
    res = run(c, empty, np.float64)
    assert res["nobs"] == 0 and not res["nobsgrd"].any()
