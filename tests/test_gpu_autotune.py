"""The in-project autotuner (letkf_autotune) and the two vector paths it
chooses between for each size bucket of the register eigensolve.

Bar: either path forced on every bucket meets the FP64 parity bar of the
analysis against the oracle (per-variable max |diff| / max |field| <= 1e-10,
local-obs counts exact); the autotuned analysis equals, bit for bit, an
analysis on a fresh handle set to the tuned paths.
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


def _problem():
    # 1-cell localisation: local-obs counts spread over every size bucket
    # (about 12 / 350 / 800 / 1500 / 170 points of size <= 16 / 32 / 48 / 64 / primal)
    return synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=64, nobs=600, seed=5, alpha_rtpp=0.7,
                              land_frac=0.1, real_kind=8, sigma_h_cells=1.0)


@pytest.mark.parametrize("vpath", [1, 2], ids=["Q", "reflectors"])
def test_forced_vector_paths_match_oracle(vpath):
    _say(f"case vector path {vpath} on every bucket, K=64 batched FP64")
    pr = _problem()
    rc, a3o, a2o, nl = po.da_main(to_oracle(pr))
    assert rc == 0
    prd = pr.to("cuda")
    eng = engine.Engine(engine.make_params(prd, "fp64", "batched"))
    try:
        eng.set_tuning([vpath] * 4 + [0])
        assert eng.get_tuning() == {16: vpath, 32: vpath, 48: vpath, 64: vpath, 128: 0}
        a3, a2 = eng.da_main_device(prd)
        torch.cuda.synchronize()
        st = eng.stats()
    finally:
        eng.close()
    assert st["local_obs"] == int(nl.sum())
    e3 = per_var_err(synth._to_fortran(a3), a3o, axis_var=3)
    e2 = per_var_err(synth._to_fortran(a2), a2o, axis_var=2)
    _say(f"  err3={max(e3):.2e} err2={max(e2):.2e}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


def test_autotune_choice_reproduces():
    _say("case autotune K=64 batched FP64")
    prd = _problem().to("cuda")
    prm = engine.make_params(prd, "fp64", "batched")
    eng = engine.Engine(prm)
    try:
        a3, a2, times, choice = eng.autotune(prd)
        torch.cuda.synchronize()
    finally:
        eng.close()
    _say(f"  times {times} choice {choice}")
    assert set(times) == {16, 32, 48, 64} and all(choice[nb] in (1, 2) for nb in times)
    for nb, (tq, tr) in times.items():
        dflt = 2 if nb >= 64 else 1
        t_d, t_a = (tr, tq) if dflt == 2 else (tq, tr)
        assert choice[nb] == ((3 - dflt) if t_a < 0.97 * t_d else dflt)
    fresh = engine.Engine(prm)
    try:
        fresh.set_tuning(choice)
        b3, b2 = fresh.da_main_device(prd)
        torch.cuda.synchronize()
    finally:
        fresh.close()
    assert torch.equal(a3, b3) and torch.equal(a2, b2)
