"""Synthetic inputs for the observation operator (Trans_XtoY,
letkf/run/src/common.f90:391-500): a global grid with land cells and
uneven bottoms, ensemble fields, and obs that reach every branch of the
operator (out of the grid, land corner, above the surface, below the bottom,
the surface level, exact level hits, the Akima interior, SSH and an
element the operator does not know)."""
import numpy as np

ID_U, ID_V, ID_T, ID_S, ID_Z = 2819, 2820, 3073, 3332, 2567


def make_grid(rng, nlon, nlat, nlev, region=None, dtype=np.float64):
    """Fortran-layout arrays as numpy arrays with reversed axes:
    lon/lat/fsm [nlat, nlon], depth [nlev, nlat, nlon].  With `region`
    ((i0, i1), (j0, j1), 1-based inclusive) only that window is filled (the
    reference's fixed 1442 x 562 x 75 grid is mostly left at zero)."""
    (i0, i1), (j0, j1) = region if region else ((1, nlon), (1, nlat))
    lon = np.zeros((nlat, nlon), dtype)
    lat = np.zeros((nlat, nlon), dtype)
    fsm = np.zeros((nlat, nlon), dtype)
    depth = np.zeros((nlev, nlat, nlon), dtype)
    I = np.arange(i0, i1 + 1)
    J = np.arange(j0, j1 + 1)
    lon[J[:, None] - 1, I[None, :] - 1] = 120.0 + (I[None, :] - 1) * 0.25
    lat[J[:, None] - 1, I[None, :] - 1] = 10.0 + (J[:, None] - 1) * 0.25
    land = rng.uniform(size=(len(J), len(I))) < 0.06
    fsm[J[:, None] - 1, I[None, :] - 1] = np.where(land, 0.0, 1.0)
    bottom = rng.uniform(50.0, 400.0, size=(len(J), len(I)))
    # levels 2..nlev from the bottom to -0.5 m, uneven spacing; level 1 a dummy below
    s = np.sort(rng.uniform(0.0, 1.0, size=(len(J), len(I), nlev - 1)), axis=2)
    s[..., 0], s[..., -1] = 0.0, 1.0
    z = -bottom[..., None] + (bottom[..., None] - 0.5) * (s ** 1.5)
    z = np.maximum.accumulate(z + 1e-3 * np.arange(nlev - 1), axis=2)
    depth[1:, J[:, None] - 1, I[None, :] - 1] = np.moveaxis(z, 2, 0)
    depth[0, J[:, None] - 1, I[None, :] - 1] = -bottom - 10.0
    return lon, lat, fsm, depth


def make_fields(rng, nmem, nv3d, nv2d, nlev, nlat, nlon, region=None, dtype=np.float64):
    """v3d [nmem, nv3d, nlev, nlat, nlon], v2d [nmem, nv2d, nlat, nlon]."""
    (i0, i1), (j0, j1) = region if region else ((1, nlon), (1, nlat))
    v3d = np.zeros((nmem, nv3d, nlev, nlat, nlon), dtype)
    v2d = np.zeros((nmem, nv2d, nlat, nlon), dtype)
    sl = (slice(j0 - 1, j1), slice(i0 - 1, i1))
    shp3 = (nmem, nv3d, nlev, j1 - j0 + 1, i1 - i0 + 1)
    base = np.cumsum(rng.standard_normal(shp3), axis=2) * 0.3 + 15.0
    v3d[:, :, :, sl[0], sl[1]] = base
    v2d[:, :, sl[0], sl[1]] = rng.standard_normal((nmem, nv2d, j1 - j0 + 1, i1 - i0 + 1)) * 0.1
    return v3d, v2d


def make_obs(rng, nobs, lon, lat, depth, region, nlon, nlat, dtype=np.float64):
    (i0, i1), (j0, j1) = region
    nlev = depth.shape[0]
    idx = rng.integers(i0, i1, size=nobs)          # i, i+1 inside the window
    idy = rng.integers(j0, j1, size=nobs)
    elm = rng.choice([ID_U, ID_V, ID_T, ID_S, ID_Z], size=nobs, p=[0.2, 0.2, 0.25, 0.2, 0.15])
    u = rng.uniform(size=nobs)
    elm[u < 0.01] = 9999                             # unknown element
    # a few indices outside [1, nlon-1] x [1, nlat-1]
    bad = rng.uniform(size=nobs) < 0.02
    idx[bad] = rng.choice([0, nlon, nlon + 3], size=bad.sum())
    bad = rng.uniform(size=nobs) < 0.02
    idy[bad] = rng.choice([0, nlat, -1], size=bad.sum())
    ii = np.clip(idx, 1, nlon - 1) - 1
    jj = np.clip(idy, 1, nlat - 1) - 1
    lono = lon[jj, ii] + rng.uniform(0, 0.25, nobs)
    lato = lat[jj, ii] + rng.uniform(0, 0.25, nobs)
    levo = np.empty(nobs)
    kind = rng.uniform(size=nobs)
    top = depth[nlev - 1, jj, ii]
    bot = depth[1, jj, ii]
    levo[:] = rng.uniform(0, 1, nobs) * (bot - top) + top                 # interior
    m = kind < 0.10
    levo[m] = rng.uniform(top[m], 0.0)                                    # between top level and 0
    m = (kind >= 0.10) & (kind < 0.18)
    k = rng.integers(1, nlev, size=nobs)
    levo[m] = depth[k[m], jj[m], ii[m]]                                   # exact level hit
    m = (kind >= 0.18) & (kind < 0.22)
    levo[m] = rng.uniform(0.1, 5.0, m.sum())                              # above the surface
    m = (kind >= 0.22) & (kind < 0.26)
    levo[m] = bot[m] - rng.uniform(0.1, 50.0, m.sum())                    # below the bottom
    m = (kind >= 0.26) & (kind < 0.30)
    levo[m] = 0.0
    return (elm.astype(np.int32), idx.astype(np.int32), idy.astype(np.int32),
            lono.astype(dtype), lato.astype(dtype), levo.astype(dtype))
