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
