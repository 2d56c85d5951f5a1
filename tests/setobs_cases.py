"""Synthetic inputs for set_obs after the observation operator
(letkf/run/src/common_obs.f90:139-402): unsorted obs of every element with
surface and subsurface levels, ensemble H(x) around the obs values, a qc
vector with rejected obs, cells outside the grid, several obs per cell and
time slots.  Every obs gets a unique longitude, so a sorted table tells which
input obs each row came from."""
import numpy as np

ID_U, ID_V, ID_T, ID_S, ID_Z = 2819, 2820, 3073, 3332, 2567
UNKNOWN = 9999


def make_case(rng, nobs, nmem, nlon, nlat, dtype=np.float64, p_qc0=0.08, p_out=0.03, crowd=0.3,
              nslots=1):
    elm = rng.choice([ID_U, ID_V, ID_T, ID_S, ID_Z, UNKNOWN], size=nobs,
                     p=[0.18, 0.18, 0.25, 0.2, 0.15, 0.04]).astype(np.int32)
    lev = np.where(rng.uniform(size=nobs) < 0.4, 0.0, -rng.uniform(1.0, 500.0, size=nobs))
    lev[elm == ID_Z] = 0.0
    idx = rng.integers(1, nlon, size=nobs).astype(np.int32)   # 1..nlon-1
    idy = rng.integers(1, nlat, size=nobs).astype(np.int32)   # 1..nlat-1
    # crowd a few cells so that cells hold many obs whose order matters
    hot = rng.uniform(size=nobs) < crowd
    idx[hot] = rng.integers(1, min(4, nlon - 1) + 1, size=hot.sum())
    idy[hot] = rng.integers(1, min(3, nlat - 1) + 1, size=hot.sum())
    out = rng.uniform(size=nobs) < p_out
    bad = rng.choice([0, 999, nlon, -3], size=nobs)
    bad_y = rng.choice([0, 999, nlat], size=nobs)
    which = rng.uniform(size=nobs) < 0.5
    idx[out & which] = bad[out & which]
    idy[out & ~which] = bad_y[out & ~which]
    lon = (100.0 + np.arange(nobs) * 1e-3)[rng.permutation(nobs)]
    lat = rng.uniform(-60.0, 60.0, size=nobs)
    dat = rng.standard_normal(nobs) * 3.0 + 15.0
    err = rng.uniform(0.05, 1.5, size=nobs)
    # H(x): obs value + an offset (the innovation) + member spread
    off = rng.standard_normal(nobs) * rng.choice([0.3, 1.5, 6.0], size=nobs)
    hx = (dat - off)[:, None] + rng.standard_normal((nobs, nmem)) * rng.uniform(0.05, 2.0, size=(nobs, 1))
    qc = (rng.uniform(size=nobs) >= p_qc0).astype(np.int32)
    hx[qc == 0, rng.integers(0, nmem)] = 999.0
    ins = rng.integers(1, 6, size=nobs).astype(np.int32)
    slots = None
    if nslots > 1:
        cut = np.sort(rng.integers(0, nobs + 1, size=nslots - 1))
        slots = np.diff(np.concatenate([[0], cut, [nobs]])).astype(np.int32)
    d = np.dtype(dtype)
    return dict(elm=elm, lon=lon.astype(d), lat=lat.astype(d), lev=lev.astype(d), dat=dat.astype(d),
                err=err.astype(d), idx=idx, idy=idy, hx=hx.astype(d), qc=qc, ins=ins, nobslots=slots)
