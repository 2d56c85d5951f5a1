"""Shared helpers of the test suite: problem conversion for the oracle and
error metrics."""
import numpy as np

import pyoracle as po


def to_oracle(pr):
    """synth.Problem (numpy form) -> pyoracle.Problem."""
    pn = pr.numpy() if not isinstance(pr.fcst3d, np.ndarray) else pr
    kw = {k: getattr(pn, k) for k in ("lon1", "lat1", "ri1", "rj1", "dlon1", "dlat1", "phi1",
                                       "depth1", "nobsgrd", "obselm", "obslon", "obslat",
                                       "obslev", "obserr", "obsdep", "obshdxf", "fcst3d", "fcst2d")}
    return po.Problem(params=dict(pn.params), **kw)


def rel_err(a, b):
    """max |a-b| / max |b| over the array (scale-aware relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def per_var_err(a, b, axis_var=-1):
    """Per-variable max |a-b| / sd(b) (field-scale relative error)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    out = []
    for v in range(a.shape[axis_var]):
        av = np.take(a, v, axis=axis_var)
        bv = np.take(b, v, axis=axis_var)
        sc = max(np.abs(bv).max(), 1e-300)
        out.append(float(np.abs(av - bv).max() / sc))
    return out


def oracle_tables(prd):
    """Host oracle problem holding the grid and observation tables of the
    device problem prd, without its ensemble (for analyse_points)."""
    import torch
    from sbpom_letkf_v2_amd import synth
    m = prd.meta
    K, nv3, nv2 = m["nmem"], m["nv3d"], m["nv2d"]
    tabs = synth.Problem(
        params=dict(prd.params), **{k: getattr(prd, k) for k in synth.Problem.ARRAYS
                                    if k not in ("fcst3d", "fcst2d")},
        fcst3d=torch.zeros((nv3, K, 1, 1), dtype=torch.float64),
        fcst2d=torch.zeros((max(nv2, 1), K, 1), dtype=torch.float64)).numpy()
    return po.Problem(params=dict(tabs.params), **{k: getattr(tabs, k) for k in synth.Problem.ARRAYS})


def sample_vs_oracle(prd, opr, anal, pts, dtype=np.float64, nthreads=0):
    """The oracle's analysis at the flat points pts (p = i + nij1*k) of the
    device problem prd, next to the device analysis anal = (anal3d, anal2d)
    at the same points.  Returns (oracle xa3, oracle xa2, oracle nobsl,
    device xa3, device xa2, surface mask), xa3 (npt, nv3d, nmem) and xa2
    (npt, nv2d, nmem)."""
    import torch
    m = prd.meta
    K, nv3, nv2 = m["nmem"], m["nv3d"], m["nv2d"]
    nij1 = m["nlon"] * m["nlat"]
    P = nij1 * m["nlev"]
    pts = np.sort(np.asarray(pts, np.int64))
    pk = (pts // nij1).astype(np.int32)
    pi = (pts % nij1).astype(np.int32)
    dev = prd.fcst3d.device
    tp = torch.as_tensor(pts, device=dev)
    ti = torch.as_tensor(pi.astype(np.int64), device=dev)

    def at(a3, a2):
        # a3 memory [v][m][p], a2 [v][m][i]
        x3 = a3.reshape(nv3, K, P)[:, :, tp].permute(2, 0, 1).contiguous().cpu().numpy()
        x2 = (a2.reshape(nv2, K, nij1)[:, :, ti].permute(2, 0, 1).contiguous().cpu().numpy()
              if nv2 else np.zeros((len(pts), 0, K)))
        return x3, x2

    f3, f2 = at(prd.fcst3d, prd.fcst2d)
    g3, g2 = at(*anal)
    rc, o3, o2, nl = po.analyse_points(opr, pi, pk, f3.astype(dtype), f2.astype(dtype), dtype=dtype,
                                       nthreads=nthreads)
    assert rc == 0, rc
    return o3, o2, nl, g3, g2, pk == m["nlev"] - 1


def mean_spread_errors(a3, a2, r3, r2, nv3, nv2, K):
    """Per-variable max error of the analysis ensemble mean and spread of
    (a3, a2) against (r3, r2) (device tensors in the engine layout), relative
    to the max |reference mean| and the max reference spread of that
    variable; also the member-level max error relative to max |reference|."""
    out = []
    for v in range(nv3 + nv2):
        if v < nv3:
            x, r = a3.reshape(nv3, K, -1)[v], r3.reshape(nv3, K, -1)[v]
        else:
            x, r = a2.reshape(nv2, K, -1)[v - nv3], r2.reshape(nv2, K, -1)[v - nv3]
        xm, rm = x.mean(0), r.mean(0)
        xs, rs = x.std(0), r.std(0)
        out.append(dict(mean=float((xm - rm).abs().max() / rm.abs().max()),
                        sprd=float((xs - rs).abs().max() / rs.max()),
                        member=float((x - r).abs().max() / r.abs().max())))
    return out
