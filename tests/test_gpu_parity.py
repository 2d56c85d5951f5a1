"""GPU parity: the HIP engine through the C ABI against the CPU oracle on the
same seeded inputs.

Tolerances (stated per test):
  * FP64 engine vs FP64 oracle: per-variable max |diff| / max |field|
    <= 1e-10 (both are FP64 LETKF; only summation order and the
    re-associated weight formulas differ).
  * REAL(4) I/O: engine vs the FP32-build oracle <= 2e-5 of the field
    scale (the FP32 reference itself carries ~1e-6 rounding noise).
Local-obs counts must agree exactly.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from helpers import per_var_err, to_oracle
from sbpom_letkf_v2_amd import engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run_both(pr, precision="fp64", form="plain", device=True):
    opr = to_oracle(pr)
    dt = np.float32 if pr.params["real_kind"] == 4 else np.float64
    rc, a3o, a2o, nl = po.da_main(opr, dtype=dt)
    assert rc == 0
    if device:
        prd = pr.to("cuda")
        eng = engine.Engine(engine.make_params(prd, precision, form))
        a3, a2 = eng.da_main_device(prd)
        torch.cuda.synchronize()
        st = eng.stats()
        eng.close()
        a3 = synth._to_fortran(a3)
        a2 = synth._to_fortran(a2)
    else:
        eng = engine.Engine(engine.make_params(pr, precision, form))
        a3, a2 = eng.da_main_host(pr.numpy())
        st = eng.stats()
        eng.close()
    return a3, a2, a3o, a2o, nl, st


@pytest.mark.parametrize("K,nobs,seed", [(20, 2000, 1), (32, 800, 2), (64, 1500, 3)])
def test_fp64_small(K, nobs, seed):
    _say(f"case fp64 K={K} nobs={nobs} seed={seed}")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=K, nobs=nobs, seed=seed,
                            alpha_rtpp=0.7, land_frac=0.1, real_kind=8)
    a3, a2, a3o, a2o, nl, st = _run_both(pr)
    assert st["local_obs"] == int(nl.sum())
    assert st["active_points"] == int((nl > 0).sum())
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


@pytest.mark.parametrize("K", [100, 128])
def test_fp64_large_ensemble(K):
    _say(f"case fp64 K={K}")
    pr = synth.make_problem(nlon=16, nlat=16, nlev=4, nmem=K, nobs=900, seed=K,
                            alpha_rtpp=0.9, real_kind=8)
    a3, a2, a3o, a2o, nl, st = _run_both(pr)
    e3 = per_var_err(a3, a3o, axis_var=3)
    _say(f"  err3={e3} maxn={nl.max()}")
    assert max(e3) <= 1e-10


def test_fp64_rtps_host_entry():
    _say("case fp64 RTPS+RTPP host entry")
    pr = synth.make_problem(nlon=20, nlat=18, nlev=5, nmem=24, nobs=1200, seed=7,
                            alpha_rtpp=0.5, alpha_rtps=0.6, real_kind=8)
    a3, a2, a3o, a2o, nl, st = _run_both(pr, device=False)
    assert max(per_var_err(a3, a3o, axis_var=3)) <= 1e-10
    assert max(per_var_err(a2, a2o, axis_var=2)) <= 1e-10


def test_fp32_io():
    _say("case real_kind=4")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=20, nobs=2000, seed=11,
                            alpha_rtpp=0.7, real_kind=4)
    a3, a2, a3o, a2o, nl, st = _run_both(pr)
    assert st["local_obs"] == int(nl.sum())
    e3 = per_var_err(a3, a3o, axis_var=3)
    _say(f"  err3={e3}")
    assert max(e3) <= 2e-5


def test_nobs_zero_copies_forecast():
    _say("case nobs=0")
    pr = synth.make_problem(nlon=12, nlat=10, nlev=3, nmem=8, nobs=0, seed=5, real_kind=8)
    prd = pr.to("cuda")
    a3, a2 = engine.da_main(prd)
    assert torch.equal(a3, prd.fcst3d) and torch.equal(a2, prd.fcst2d)
