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


@pytest.mark.parametrize("form", ["batched", "batched_ql"])
@pytest.mark.parametrize("K,nobs,seed", [(20, 300, 51), (20, 100, 54), (64, 700, 52),
                                         (100, 1800, 53), (128, 2000, 55)])
def test_batched_form_matches_fp64_oracle(K, nobs, seed, form):
    """Batched form: points with nobsl < K solve the nobsl x nobsl problem;
    "batched" applies the inverse and inverse square root of the
    tridiagonalised matrix by shifted solves (Zolotarev's rational function,
    relative error < 1e-15), "batched_ql" through its QL eigen-
    decomposition.  Same LETKF, so the FP64 tolerance is unchanged (1e-10).
    A 1-cell / 1-level localisation scale gives a mix of nobsl < K and
    nobsl >= K."""
    _say(f"case {form} K={K} nobs={nobs} seed={seed}")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=K, nobs=nobs, seed=seed,
                            alpha_rtpp=0.7, alpha_rtps=0.3, land_frac=0.1, real_kind=8,
                            sigma_h_cells=1.0, sigma_v=1.0)
    a3, a2, a3o, a2o, nl, st = _run_both(pr, form=form)
    act = nl[nl > 0]
    _say(f"  n<K points {(act < K).sum()} n>=K points {(act >= K).sum()} fallback {st['fallback_slots']}")
    assert (act < K).sum() > 0 and st["local_obs"] == int(nl.sum())
    assert st["fallback_slots"] == 0
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


@pytest.mark.parametrize("K,scale", [(32, 0.01), (64, 0.01), (100, 0.01), (64, 0.003)])
def test_batched_rational_ill_conditioned(K, scale):
    """Obs errors shrunk by `scale`: the local matrices' spectra widen to
    kappa ~ 1e4-1e7 (up to ~36 poles of the rational function); the FP64
    batched form stays within 1e-10 of the oracle, with no fallback at
    scale 0.01; at 0.003 the widest spectra pass the table's last class
    (2^24) and those slots take the fallback within the same analysis."""
    _say(f"case batched ill-conditioned K={K} obs error x{scale}")
    pr = synth.make_problem(nlon=20, nlat=18, nlev=5, nmem=K, nobs=1500, seed=60 + K,
                            alpha_rtpp=0.5, land_frac=0.1, real_kind=8)
    pr.obserr = pr.obserr * scale
    a3, a2, a3o, a2o, nl, st = _run_both(pr, form="batched")
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  maxn={nl.max()} fallback {st['fallback_slots']} err3={max(e3):.2e} err2={max(e2):.2e}")
    if scale >= 0.01:
        assert st["fallback_slots"] == 0
    else:
        assert 0 < st["fallback_slots"] < st["active_points"]
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


@pytest.mark.parametrize("K", [48, 100])
def test_batched_rational_fallback(K, monkeypatch):
    """LETKF_RAT_CLASSES=0 (a test hook) removes every class of the rational
    table, so every active slot takes the fallback: the LDS eigensolve and
    the tile update with the reference's eigenvalue rule."""
    _say(f"case batched forced fallback K={K}")
    monkeypatch.setenv("LETKF_RAT_CLASSES", "0")
    pr = synth.make_problem(nlon=16, nlat=14, nlev=4, nmem=K, nobs=700, seed=70 + K,
                            alpha_rtpp=0.7, alpha_rtps=0.2, real_kind=8, sigma_h_cells=1.5)
    a3, a2, a3o, a2o, nl, st = _run_both(pr, form="batched")
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  active {st['active_points']} fallback {st['fallback_slots']} err3={max(e3):.2e} err2={max(e2):.2e}")
    assert st["fallback_slots"] == st["active_points"] > 0
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


def _csr_problems(rng, K, sizes, dtype):
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    L = int(offs[-1])
    Y = rng.standard_normal((max(L, 1), K))
    for b in range(len(sizes)):
        seg = slice(offs[b], offs[b + 1])
        Y[seg] -= Y[seg].mean(1, keepdims=True)
    rd = rng.uniform(0.3, 2.0, max(L, 1)) ** 2
    rl = rng.uniform(0.01, 1.0, max(L, 1))
    dp = rng.standard_normal(max(L, 1))
    return offs, [np.asarray(a, dtype) for a in (Y, rd, rl, dp)]


@pytest.mark.parametrize("K", [20, 64, 100, 128])
def test_core_batched_vs_oracle(K):
    """letkf_core_batched (wvec, Wmat of common_letkf.f90:37-210) vs the
    FP64 oracle's letkf_core per problem, including empty problems."""
    import ctypes as C
    from sbpom_letkf_v2_amd import _lib
    _say(f"case core_batched K={K}")
    rng = np.random.default_rng(K)
    sizes = [0, 1, 3, K - 1, K, 2 * K + 5, 17, 0, 250]
    offs, (Y, rd, rl, dp) = _csr_problems(rng, K, sizes, np.float64)
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=K, nobs=10, seed=1)
    prm = engine.make_params(pr)
    prm.cov_infl_mul = 1.15
    eng = engine.Engine(prm)
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    tO, tY, trd, trl, tdp = dev(offs), dev(Y), dev(rd), dev(rl), dev(dp)
    nb = len(sizes)
    w = torch.empty((nb, K), dtype=torch.float64, device="cuda")
    W = torch.empty((nb, K, K), dtype=torch.float64, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    _lib.check(_lib.lib().letkf_core_batched(eng.h, nb, vp(tO), vp(tY), vp(trd), vp(trl), vp(tdp),
                                             vp(w), vp(W)))
    w, W = w.cpu().numpy(), W.cpu().numpy()
    eng.close()
    for b, n in enumerate(sizes):
        seg = slice(offs[b], offs[b + 1])
        if n == 0:
            assert np.allclose(W[b], np.sqrt(1.15) * np.eye(K), atol=1e-15) and not w[b].any()
            continue
        rc, wo, Wo = po.letkf_core(Y[seg], rd[seg], rl[seg], dp[seg], 1.15)
        assert np.abs(w[b] - wo).max() <= 1e-10 * max(np.abs(wo).max(), 1e-30)
        assert np.abs(W[b] - Wo).max() <= 1e-10 * np.abs(Wo).max()   # wmat[b][i][j] = Wmat(i,j)


def test_bin_obs_bit_exact_vs_oracle():
    import ctypes as C
    from sbpom_letkf_v2_amd import _lib
    _say("case bin_obs")
    rng = np.random.default_rng(77)
    nlon, nlat = 37, 23
    lon = 130 + np.arange(nlon) / 16
    lat = 10 + np.arange(nlat) / 16
    L2 = np.asfortranarray(np.repeat(lon[:, None], nlat, 1))
    A2 = np.asfortranarray(np.repeat(lat[None, :], nlon, 0))
    n = 5000
    olon = rng.uniform(lon[0] - 0.05, lon[-1] + 0.05, n)
    olat = rng.uniform(lat[0] - 0.05, lat[-1] + 0.05, n)
    olon[:100] = lon[rng.integers(0, nlon, 100)]
    idx_o, idy_o, perm_o, grd_o = po.bin_obs(L2, A2, olon, olat)
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=4, nobs=10, seed=1)
    eng = engine.Engine(engine.make_params(pr))
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    tL, tA = dev(L2.T), dev(A2.T)            # memory = Fortran (nlon,nlat)
    to, ta = dev(olon), dev(olat)
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    idy = torch.empty_like(idx)
    perm = torch.empty_like(idx)
    grd = torch.empty((nlat, nlon), dtype=torch.int32, device="cuda")
    ns = C.c_int32()
    vp = lambda t: C.c_void_p(t.data_ptr())
    _lib.check(_lib.lib().letkf_bin_obs(eng.h, nlon, nlat, vp(tL), vp(tA), n, vp(to), vp(ta), vp(idx),
                                        vp(idy), vp(perm), vp(grd), C.byref(ns)))
    eng.close()
    assert ns.value == len(perm_o)
    assert np.array_equal(idx.cpu().numpy(), idx_o) and np.array_equal(idy.cpu().numpy(), idy_o)
    assert np.array_equal(perm.cpu().numpy()[: ns.value], perm_o)
    assert np.array_equal(grd.cpu().numpy().T, grd_o)


def test_aoei_gpu_vs_oracle():
    import ctypes as C
    from sbpom_letkf_v2_amd import _lib
    _say("case aoei")
    rng = np.random.default_rng(3)
    e = rng.uniform(0.1, 1, 1000)
    d = rng.standard_normal(1000)
    s = rng.uniform(0, 1, 1000)
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=4, nobs=10, seed=1)
    eng = engine.Engine(engine.make_params(pr))
    te, td, ts = (torch.as_tensor(a).cuda() for a in (e, d, s))
    vp = lambda t: C.c_void_p(t.data_ptr())
    _lib.check(_lib.lib().letkf_aoei(eng.h, 1000, vp(te), vp(td), vp(ts)))
    eng.close()
    assert np.array_equal(te.cpu().numpy(), po.aoei(e, d, s))


def test_rotation_log_overflow_fallback(monkeypatch):
    """Slots whose QL rotation log overflows are solved whole by the LDS
    eigensolver and updated by the full tile transform.  A capped log
    (LETKF_ROT_LOG_CAP) sends the larger slots down that path while the
    small ones stay on the register path; both must meet the FP64 bar."""
    monkeypatch.setenv("LETKF_ROT_LOG_CAP", "300")
    _say("case rotation-log overflow fallback K=32")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=32, nobs=800, seed=71, alpha_rtpp=0.7,
                            alpha_rtps=0.3, land_frac=0.1, real_kind=8, sigma_h_cells=1.0, sigma_v=1.0)
    a3, a2, a3o, a2o, nl, st = _run_both(pr, form="batched_ql")
    _say(f"  fallback slots {st['fallback_slots']} of {st['active_points']}")
    assert 0 < st["fallback_slots"] < st["active_points"]
    e3 = per_var_err(a3, a3o, axis_var=3)
    e2 = per_var_err(a2, a2o, axis_var=2)
    _say(f"  err3={e3} err2={e2}")
    assert max(e3) <= 1e-10 and max(e2) <= 1e-10


@pytest.mark.parametrize("precision,form,K,nobs", [("fp32", "batched", 48, 900), ("bf16x3", "batched", 48, 900),
                                                   ("fp32", "plain", 48, 900), ("fp32", "batched", 100, 1800),
                                                   ("bf16x3", "batched", 128, 2000), ("fp32", "plain", 100, 1800),
                                                   ("fp32p", "batched", 48, 900), ("fp32p", "plain", 100, 1800),
                                                   ("fp32x3", "batched", 48, 900), ("fp32x3", "batched", 100, 1800),
                                                   ("fp32d", "batched", 48, 900), ("fp32d", "batched", 100, 1800),
                                                   ("bf16x3t", "batched", 48, 900), ("bf16x3t", "batched", 100, 1800),
                                                   ("bf16x3t", "batched", 128, 2000)])
def test_mixed_precision_gram_close_to_fp64(precision, form, K, nobs):
    """Gram on MFMA (FP32, FP32 summed per panel, FP32 or BF16 hi/lo splits) with FP64 eigensolve
    (bf16x3t: the Householder reduction in FP32)
    against the FP64 oracle: the analysis-mean max error <= 1e-5 relative to
    the variable's max |mean| (the target bound of the mixed-precision
    form), every member within 1e-4 of the field scale (FP32 products keep
    ~7 digits; the innovation term stays FP64).  K > 64 runs the 128-wide
    output tiling, partly filled for K = 100."""
    _say(f"case {precision}/{form} K={K} nobs={nobs}")
    pr = synth.make_problem(nlon=24, nlat=20, nlev=6, nmem=K, nobs=nobs, seed=61, alpha_rtpp=0.7,
                            real_kind=8, sigma_h_cells=1.0, sigma_v=1.0)
    a3, a2, a3o, a2o, nl, st = _run_both(pr, precision=precision, form=form)
    e3 = per_var_err(a3, a3o, axis_var=3)
    m3 = per_var_err(a3.mean(2), a3o.mean(2), axis_var=2)
    m2 = per_var_err(a2.mean(1), a2o.mean(1), axis_var=1)
    _say(f"  members err3={max(e3):.2e}; mean err3={m3} err2={m2}")
    assert max(e3) <= 1e-4
    assert max(m3) <= 1e-5 and max(m2) <= 1e-5


def test_c4_cycles_mixed_k():
    """Workload c4 at reduced size: eight independent cycles, nmem round-robin
    over 32, 64, 100, 128, 0.5 .. 4 obs per column, run in sequence on fresh
    handles.  Each cycle's mixed-precision analysis (FP32 and BF16 hi/lo
    Gram alternating) stays within 1e-4 of the FP64 engine on the same
    cycle, its ensemble mean within 1e-5; the FP64 engine meets the oracle (1e-10) on the nmem <= 64 cycles."""
    cycles = synth.c4_cycles(scale=1.0 / 16.0, nlev=6)
    for c, kw in enumerate(cycles):
        prec = "fp32" if c % 2 == 0 else "bf16x3"
        _say(f"case c4 cycle {c} nlon={kw['nlon']} nmem={kw['nmem']} nobs={kw['nobs']} {prec}")
        pr = synth.make_problem(device="cuda", real_kind=8, sigma_h_cells=1.5, sigma_v=1.0, **kw)
        outs = {}
        for name, pf in (("fp64", ("fp64", "batched")), ("mixed", (prec, "batched"))):
            eng = engine.Engine(engine.make_params(pr, *pf))
            a3, a2 = eng.da_main_device(pr)
            torch.cuda.synchronize()
            eng.close()
            outs[name] = (synth._to_fortran(a3), synth._to_fortran(a2))
        e3 = per_var_err(outs["mixed"][0], outs["fp64"][0], axis_var=3)
        m3 = per_var_err(outs["mixed"][0].mean(2), outs["fp64"][0].mean(2), axis_var=2)
        _say(f"  mixed vs FP64 members err3={max(e3):.2e} mean err3={max(m3):.2e}")
        assert max(e3) <= 1e-4 and max(m3) <= 1e-5
        if kw["nmem"] <= 64:
            opr = to_oracle(pr.to("cpu"))
            rc, a3o, a2o, nl = po.da_main(opr, dtype=np.float64)
            assert rc == 0
            e = per_var_err(outs["fp64"][0], a3o, axis_var=3)
            _say(f"  FP64 vs oracle err3={max(e):.2e}")
            assert max(e) <= 1e-10

