"""The in-project autotuner (letkf_autotune) and the three vector paths it
chooses between for each size bucket of the register eigensolve (1: Q
accumulated from the Householder reflectors; 2: the reflectors applied one
wavefront sum per vector; 3: the reflectors applied in the (row group,
column) lane layout).

Bar: every path forced on every bucket meets the FP64 parity bar of the
analysis against the oracle (per-variable max |diff| / max |field| <= 1e-10,
local-obs counts exact), at K = 64 (buckets 16..64) and K = 100 (buckets
16..128, paths 2 and 3); the autotuned analysis equals, bit for bit, an
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


# the fixed rule's vector path per bucket at K = 64 with 6 vector columns
DEFAULT_PATH = {16: 3, 32: 3, 48: 3, 64: 3}


def _problem():
    # 1-cell localisation: local-obs counts spread over every size bucket
    # (about 12 / 350 / 800 / 1500 / 170 points of size <= 16 / 32 / 48 / 64 / primal)
    return synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=64, nobs=600, seed=5, alpha_rtpp=0.7,
                              land_frac=0.1, real_kind=8, sigma_h_cells=1.0)


@pytest.mark.parametrize("vpath", [1, 2, 3], ids=["Q", "reflectors", "blocked"])
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


@pytest.mark.parametrize("vpath", [2, 3], ids=["reflectors", "blocked"])
@pytest.mark.parametrize("form", ["batched", "plain"])
def test_forced_vector_paths_k100(vpath, form):
    """K = 100: two members per lane, the 128 bucket (primal K x K form and
    the dual form at 65..99 local obs), and in the plain form the QL path
    (E2 + V2) between V1 and V3."""
    _say(f"case vector path {vpath} on every bucket, K=100 {form} FP64")
    pr = synth.make_problem(nlon=20, nlat=16, nlev=5, nmem=100, nobs=600, seed=11, alpha_rtpp=0.6,
                            land_frac=0.1, real_kind=8, sigma_h_cells=1.0)
    rc, a3o, a2o, nl = po.da_main(to_oracle(pr))
    assert rc == 0
    nl = np.asarray(nl).ravel()
    for lo, hi in ((16, 32), (32, 48), (48, 64), (64, 99), (99, 1 << 30)):  # every bucket, and primal points
        assert ((nl > lo) & (nl <= hi)).any(), (lo, hi)
    prd = pr.to("cuda")
    eng = engine.Engine(engine.make_params(prd, "fp64", form))
    try:
        eng.set_tuning([vpath] * 5)
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
    assert set(times) == {16, 32, 48, 64}
    for nb, t in times.items():
        assert set(t) == {1, 2, 3} and choice[nb] in t
        dflt = DEFAULT_PATH[nb]
        faster = {v: ms for v, ms in t.items() if ms < 0.97 * t[dflt]}
        expect = min(faster, key=faster.get) if faster else dflt
        assert choice[nb] == expect
    fresh = engine.Engine(prm)
    try:
        fresh.set_tuning(choice)
        b3, b2 = fresh.da_main_device(prd)
        torch.cuda.synchronize()
    finally:
        fresh.close()
    assert torch.equal(a3, b3) and torch.equal(a2, b2)
