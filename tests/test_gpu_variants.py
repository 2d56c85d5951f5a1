"""The kernel variants the defaults leave out, against the FP64 oracle.

The column search, the one-wavefront Gram (on pre-split rows in the dual
form), the MFMA Householder reduction and the point-major copies are the
defaults; the kernels they replaced stay
in the library (the per-point search is also the path for nlev > 128) and
are reached through switches read at plan / launch time:

  LETKF_SEARCH_POINT    local-obs count and fill one thread per point;
  LETKF_GRAM_BLOCK      the BF16x3 Gram on the four-wavefront block kernel;
  LETKF_GRAM_NOSPLIT    the dual-form BF16x3 Gram splitting the FP64 rows per
                        gather instead of reading the pre-split table;
  LETKF_E1_VALU         the FP32 reduction of the 16 bucket one matrix a
                        wavefront and of the 32 / 64 buckets on the vector
                        ALUs;
  LETKF_NO_POINT_MAJOR  the register update on the caller's arrays;
  LETKF_EXACT_RLOC      rloc from the FP64 exponential in precision bf16x3t
                        (the FP32 one otherwise);
  LETKF_NO_PM_STATS     the forecast mean and spread formed by the vector
                        kernels instead of the point-major copy.

Bars: FP64 analyses <= 1e-10 of the field scale and local-obs counts equal
to the oracle's; the bf16x3t analysis-mean error <= 1e-5 (members 1e-4),
as in test_gpu_parity.
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


def _problem(K):
    # every size bucket and primal points (1-cell localisation)
    return synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=K, nobs=700 if K == 64 else 900, seed=23 + K,
                              alpha_rtpp=0.7, land_frac=0.1, real_kind=8, sigma_h_cells=1.0)


def _run(pr, precision):
    prd = pr.to("cuda")
    eng = engine.Engine(engine.make_params(prd, precision, "batched"))
    try:
        a3, a2 = eng.da_main_device(prd)
        torch.cuda.synchronize()
        st = eng.stats()
    finally:
        eng.close()
    return synth._to_fortran(a3), synth._to_fortran(a2), st


def _mean_err(a3, a3o):
    # analysis-mean error per variable, relative to the field maximum
    m, mo = a3.mean(axis=2), a3o.mean(axis=2)
    return max(float(np.abs(m[..., v] - mo[..., v]).max() / np.abs(mo[..., v]).max()) for v in range(a3.shape[3]))


@pytest.mark.parametrize("hook", ["LETKF_SEARCH_POINT", "LETKF_GRAM_BLOCK", "LETKF_GRAM_NOSPLIT", "LETKF_E1_VALU",
                                  "LETKF_NO_POINT_MAJOR", "LETKF_EXACT_RLOC", "LETKF_NO_PM_STATS", None])
@pytest.mark.parametrize("K", [64, 100])
def test_variant_matches_oracle(monkeypatch, hook, K):
    _say(f"case variant {hook} K={K}")
    if hook:
        monkeypatch.setenv(hook, "1")
    pr = _problem(K)
    rc, a3o, a2o, nl = po.da_main(to_oracle(pr))
    assert rc == 0
    a3, a2, st = _run(pr, "fp64")
    assert st["local_obs"] == int(nl.sum())
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  fp64 err3={max(e3):.2e} err2={max(e2):.2e} buckets={st['bucket_slots']}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10
    b3, b2, _ = _run(pr, "bf16x3t")
    em = _mean_err(b3, a3o)
    e3 = per_var_err(b3, a3o, axis_var=3)
    _say(f"  bf16x3t mean err={em:.2e} member err={max(e3):.2e}")
    assert em <= 1e-5 and max(e3) <= 1e-4
