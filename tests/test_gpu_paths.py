"""GPU parity of the code paths the default workloads leave out, against the
FP64 oracle (FP64: <= 1e-10 of the field scale; REAL(4): <= 2e-5):

  * round_off (common_da.f90:506-533) made to fire: T and S members pushed
    past tmin/tmax/smin/smax at random points, with land points
    (phi1 == 1) and the bottom level k = 1 that the clamp skips; through
    the register eigensolve's update, the update of points without local
    obs, the fallback (LETKF_ROT_LOG_CAP, LETKF_RAT_CLASSES), the 128 bucket, the
    LPF core and REAL(4) I/O;
  * the tracer layout (5 3-D variables + SSH: 7 vector columns) and the
    widest one (7 + 1 variables: 9 columns) through every vector path of
    the register eigensolve;
  * failure paths: NaN in obsdep, Inf in obshdxf and obserr = 0 return
    LETKF_ERR_NONFINITE with every kernel finished and the analysis of the
    points the bad value does not reach unchanged; points with a single
    local observation.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from helpers import per_var_err, to_oracle
from sbpom_letkf_v2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _oracle(pr, dtype=np.float64, lpfgm=None):
    rc, a3o, a2o, nl = po.da_main(to_oracle(pr), dtype=dtype, lpfgm=lpfgm)
    assert rc == 0
    return a3o, a2o, nl


def _engine(pr, precision="fp64", form="batched", paths=None, lpfgm=None):
    prd = pr.to("cuda")
    eng = engine.Engine(engine.make_params(prd, precision, form))
    try:
        if paths is not None:
            eng.set_tuning(paths)
        if lpfgm is None:
            a3, a2 = eng.da_main_device(prd)
        else:
            a3, a2 = eng.da_main_lpfgm_device(prd, lpfgm["rand"], mode=lpfgm["mode"])
        torch.cuda.synchronize()
        st = eng.stats()
    finally:
        eng.close()
    return synth._to_fortran(a3), synth._to_fortran(a2), st


def _push_out_of_bounds(pr, seed):
    """Shift T and S of every member at random points past the round_off
    bounds (T +45 / -25, S +20 / -12 from 15 / 34.5): the perturbations are
    unchanged, the forecast mean leaves [tmin, tmax] / [smin, smax]."""
    f = pr.fcst3d                                   # memory (nij1, nlev, nmem, nv3d)
    g = torch.Generator().manual_seed(seed)
    sel = torch.rand(f.shape[2:], generator=g, dtype=torch.float64)   # (nlev, nij1)
    one = torch.ones((), dtype=f.dtype)
    T, S = f[2], f[3]
    T[:, sel < 0.15] += 45 * one
    T[:, (sel >= 0.15) & (sel < 0.3)] -= 25 * one
    S[:, (sel >= 0.3) & (sel < 0.45)] += 20 * one
    S[:, (sel >= 0.45) & (sel < 0.6)] -= 12 * one
    return pr


def _check_clamps(a3o, pr):
    """The oracle's analysis is clamped where round_off applies and keeps
    out-of-bounds values on land and at k = 1 (so the case tests both)."""
    p = pr.params
    T, S = a3o[:, :, :, 2], a3o[:, :, :, 3]
    phi = pr.numpy().phi1
    land = phi == 1.0
    sea = ~land
    tmax, tmin, smax, smin = (np.asarray(p[k], a3o.dtype) for k in ("tmax", "tmin", "smax", "smin"))
    for X, lo, hi in ((T, tmin, tmax), (S, smin, smax)):
        assert (X[sea, 1:] == hi).any() and (X[sea, 1:] == lo).any()
        assert X[sea, 1:].max() <= hi and X[sea, 1:].min() >= lo
        assert X[:, 0].max() > hi or X[:, 0].min() < lo          # k = 1 is not clamped
        assert X[land].max() > hi or X[land].min() < lo           # nor is land


@pytest.mark.parametrize("K,nobs,form,real_kind", [(20, 8, "batched", 8), (48, 400, "plain", 8),
                                                   (20, 8, "batched", 4), (128, 300, "plain", 8)])
def test_round_off_fires(K, nobs, form, real_kind):
    """The register eigensolve's fused update, the update of points without
    local obs, the 128 bucket (K = 128), REAL(4) I/O."""
    _say(f"case round_off K={K} nobs={nobs} {form} real_kind={real_kind}")
    pr = synth.make_problem(nlon=20, nlat=16, nlev=5, nmem=K, nobs=nobs, seed=81 + K, alpha_rtpp=0.6,
                            land_frac=0.2, real_kind=real_kind, sigma_h_cells=1.0, sigma_v=1.0)
    pr = _push_out_of_bounds(pr, 5)
    dt = np.float32 if real_kind == 4 else np.float64
    a3o, a2o, nl = _oracle(pr, dt)
    _check_clamps(a3o, pr)
    _say(f"  points without local obs: {(nl == 0).sum()} of {nl.size}")
    assert (nl > 0).any()
    if nobs < 20:
        assert (nl == 0).sum() > 100
    a3, a2, st = _engine(pr, "fp64", form)
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    tol = 2e-5 if real_kind == 4 else 1e-10
    assert max(e3) <= tol and max(e2) <= tol
    if real_kind == 8:
        # clamped values are the bounds themselves
        for v, (lo, hi) in ((2, ("tmin", "tmax")), (3, ("smin", "smax"))):
            for b in (lo, hi):
                at = a3o[:, :, :, v] == pr.params[b]
                assert np.array_equal(a3[:, :, :, v][at], a3o[:, :, :, v][at])


@pytest.mark.parametrize("form,hook", [("batched_ql", ("LETKF_ROT_LOG_CAP", "300")),
                                       ("batched", ("LETKF_RAT_CLASSES", "1"))])
def test_round_off_fires_overflow_fallback(monkeypatch, form, hook):
    """Slots whose QL log overflows (batched_ql), or whose spectrum bound
    lies past the rational table's classes in use (batched: one class,
    kappa <= 2), are updated by the full tile transform."""
    monkeypatch.setenv(*hook)
    _say(f"case round_off fallback K=32 {form} {hook}")
    pr = synth.make_problem(nlon=20, nlat=16, nlev=5, nmem=32, nobs=600, seed=91, alpha_rtpp=0.6,
                            alpha_rtps=0.3, land_frac=0.2, real_kind=8, sigma_h_cells=1.5, sigma_v=1.0)
    pr = _push_out_of_bounds(pr, 6)
    a3o, a2o, nl = _oracle(pr)
    _check_clamps(a3o, pr)
    a3, a2, st = _engine(pr, "fp64", form)
    _say(f"  fallback slots {st['fallback_slots']} of {st['active_points']}")
    assert st["fallback_slots"] > 0
    e3 = per_var_err(a3, a3o, axis_var=3)
    _say(f"  err3={e3}")
    assert max(e3) <= 1e-10


def test_round_off_fires_lpf():
    _say("case round_off LPF")
    K = 24
    pr = synth.make_problem(nlon=16, nlat=12, nlev=4, nmem=K, nobs=200, seed=93, alpha_rtpp=0.6,
                            land_frac=0.2, real_kind=8, sigma_h_cells=1.5, sigma_v=1.0)
    pr = _push_out_of_bounds(pr, 7)
    rand = np.random.default_rng(4).random((20, K))
    lp = dict(mode=2, rand=rand)
    a3o, a2o, nl = _oracle(pr, lpfgm=lp)
    _check_clamps(a3o, pr)
    a3, a2, st = _engine(pr, lpfgm=lp)
    e3 = per_var_err(a3, a3o, axis_var=3)
    _say(f"  err3={e3}")
    assert max(e3) <= 1e-10


@pytest.mark.parametrize("K,nv3d,tracer,paths", [(100, 5, True, None), (100, 5, True, "B"),
                                                 (48, 5, True, "Q"), (48, 5, True, "R"), (48, 5, True, "B"),
                                                 (64, 7, True, "Q"), (64, 7, True, "R"), (64, 7, True, "B"),
                                                 (100, 7, True, None)])
def test_vector_columns_tracer_and_widest(K, nv3d, tracer, paths):
    """7 vector columns (5 3-D variables with the tracer + SSH) and 9 (7 + 1),
    K = 100 on the fixed rule's path and on the blocked reflectors, and
    K <= 64 with each path forced ("Q": Q accumulated, "R": reflectors one
    wavefront sum per column, "B": the blocked layout; 9 columns exceed
    the blocked layout's 8 and fall back to the rule)."""
    _say(f"case columns K={K} nv3d={nv3d} paths={paths}")
    pr = synth.make_problem(nlon=20, nlat=16, nlev=5, nmem=K, nv3d=nv3d, nv2d=1, tracer=tracer,
                            nobs=1500 if K == 100 else 800, seed=101 + K + nv3d, alpha_rtpp=0.7,
                            alpha_rtps=0.2, real_kind=8, sigma_h_cells=1.0, sigma_v=1.0)
    a3o, a2o, nl = _oracle(pr)
    vp = None if paths is None else {nb: {"Q": 1, "R": 2, "B": 3}[paths] for nb in (16, 32, 48, 64, 128)}
    a3, a2, st = _engine(pr, "fp64", "batched", paths=vp)
    act = nl[nl > 0]
    _say(f"  nobsl<K {(act < K).sum()} nobsl>=K {(act >= K).sum()}")
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


def _analyse_bad(pr, precision="fp64", form="batched"):
    """The analysis of a problem with a non-finite or zero input: the
    engine's status (an EngineError code or 0) and its outputs."""
    prd = pr.to("cuda")
    eng = engine.Engine(engine.make_params(prd, precision, form))
    a3 = torch.full_like(prd.fcst3d, -7.0)
    a2 = torch.full_like(prd.fcst2d, -7.0)
    code = 0
    try:
        eng.da_main_device(prd, a3, a2)
    except _lib.EngineError as e:
        code = e.code
        _say(f"  engine: {e}")
    torch.cuda.synchronize()
    eng.close()
    return code, synth._to_fortran(a3), synth._to_fortran(a2)


def _unaffected_points_agree(pr_clean, j, a3):
    """At every point whose local observations exclude obs j, the engine's
    analysis of the corrupted problem is the oracle's analysis of the clean
    one (<= 1e-10 of the field scale)."""
    opr = to_oracle(pr_clean)
    a3o, a2o, nl = _oracle(pr_clean)
    nij1, nlev = nl.shape
    hit = np.zeros((nij1, nlev), bool)
    for k in range(nlev):
        for i in range(nij1):
            if nl[i, k]:
                hit[i, k] = j in po.obs_local(opr, i, k)[0]
    ok = ~hit
    assert ok.any() and hit.any()
    err = float(np.abs(a3[ok] - a3o[ok]).max() / np.abs(a3o[ok]).max())
    _say(f"  {hit.sum()} of {hit.size} points use obs {j}; engine err at the others {err:.2e}")
    assert err <= 1e-10


@pytest.mark.parametrize("form", ["batched", "batched_ql", "plain"])
@pytest.mark.parametrize("bad", ["nan_dep", "inf_hdxf", "zero_err"])
def test_nonfinite_inputs_are_reported(bad, form):
    _say(f"case failure path {bad} {form}")
    kw = dict(nlon=20, nlat=16, nlev=4, nmem=24, nobs=300, seed=111, alpha_rtpp=0.6, real_kind=8,
              sigma_h_cells=1.5, sigma_v=1.0)
    pr = synth.make_problem(**kw)
    j = 17
    if bad == "nan_dep":
        pr.obsdep[j] = float("nan")
    elif bad == "inf_hdxf":
        pr.obshdxf[3, j] = float("inf")        # memory (nmem, nobs): member 4 of obs j
    else:
        pr.obserr[j] = 0.0
    code, a3, a2 = _analyse_bad(pr, "fp64", form)
    assert code == -5                          # LETKF_ERR_NONFINITE
    _unaffected_points_agree(synth.make_problem(**kw), j, a3)


def test_nonfinite_reported_by_mixed_precision_gram():
    _say("case failure path nan_dep fp32 Gram")
    pr = synth.make_problem(nlon=20, nlat=16, nlev=4, nmem=24, nobs=300, seed=112, alpha_rtpp=0.6,
                            real_kind=8, sigma_h_cells=1.5, sigma_v=1.0)
    pr.obshdxf[5, 40] = float("nan")
    code, a3, a2 = _analyse_bad(pr, "fp32", "batched")
    assert code == -5


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_single_local_observation_points(form):
    """Three observations far apart: most points with obs see exactly one
    (the 1 x 1 dual problem / a rank-one K x K Gram)."""
    _say(f"case single-obs points {form}")
    pr = synth.make_problem(nlon=40, nlat=30, nlev=4, nmem=20, nobs=3, seed=121, alpha_rtpp=0.6,
                            real_kind=8, sigma_h_cells=1.5, sigma_v=1.0)
    a3o, a2o, nl = _oracle(pr)
    _say(f"  points with 1 local obs: {(nl == 1).sum()}, with more: {(nl > 1).sum()}")
    assert (nl == 1).sum() > 10
    a3, a2, st = _engine(pr, "fp64", form)
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "bf16x3t"])
def test_degenerate_obs_ensemble(precision):
    """Obs whose ensemble perturbations are zero or 1e-20 of the others:
    Gram entries at and below the FP32 denormal range, columns the FP32
    Householder steps must take as zero (a reflector skipped), against the
    FP64 oracle."""
    _say(f"case degenerate obs ensemble {precision}")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=5, nmem=40, nobs=600, seed=131, alpha_rtpp=0.5,
                            real_kind=8, sigma_h_cells=1.5)
    pr.obshdxf[:, 0::3] *= 1e-20   # (nmem, nobs): the Fortran (nobs, nmem) array
    pr.obshdxf[:, 1::3] = 0.0
    a3o, a2o, nl = _oracle(pr)
    a3, a2, st = _engine(pr, precision, "batched")
    assert st["local_obs"] == int(nl.sum())
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={max(e3):.2e} err2={max(e2):.2e} buckets={st['bucket_slots']}")
    if precision == "fp64":
        assert max(e3) <= 1e-10 and max(e2) <= 1e-10
    else:
        assert max(e3) <= 1e-4 and max(e2) <= 1e-4
        m, mo = a3.mean(axis=2), a3o.mean(axis=2)
        em = max(float(np.abs(m[..., v] - mo[..., v]).max() / np.abs(mo[..., v]).max()) for v in range(a3.shape[3]))
        assert em <= 1e-5


@pytest.mark.parametrize("scale,kmax,bar", [(300.0, None, 1e-7), (10.0, "256", 1e-5)])
def test_wide_spectra_promoted_to_fp64(monkeypatch, scale, kmax, bar):
    """A third of the obs with perturbations `scale` times the others: the
    local matrices' spectra widen past what the BF16x3 Gram resolves.  Slots
    whose spectrum bound exceeds the limit (LETKF_KAPPA_MAX, default 1024 for
    bf16x3t) take the fallback on an FP64 Gram, and the analysis stays within
    `bar` of the FP64 oracle (analysis mean, relative to the field scale)."""
    _say(f"case wide spectra scale={scale} kmax={kmax}")
    if kmax:
        monkeypatch.setenv("LETKF_KAPPA_MAX", kmax)
    pr = synth.make_problem(nlon=24, nlat=20, nlev=5, nmem=40, nobs=600, seed=131, alpha_rtpp=0.5,
                            real_kind=8, sigma_h_cells=1.5)
    pr.obshdxf[:, 0::3] *= scale   # (nmem, nobs): the Fortran (nobs, nmem) array
    a3o, a2o, nl = _oracle(pr)
    a3, a2, st = _engine(pr, "bf16x3t", "batched")
    m, mo = a3.mean(axis=2), a3o.mean(axis=2)
    em = max(float(np.abs(m[..., v] - mo[..., v]).max() / np.abs(mo[..., v]).max()) for v in range(a3.shape[3]))
    _say(f"  mean err={em:.2e} fallback slots={st['fallback_slots']}")
    assert st["fallback_slots"] > 0
    assert em <= bar
