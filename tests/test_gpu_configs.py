"""Parity on the named workloads at their full sizes (synth.config).

  c0  64x64x16, K=20, 2000 random obs: every one of the 65,536 points
      against the oracle -- FP64 engine (plain and batched forms) vs the
      FP64 oracle <= 1e-10 of the field scale, REAL(4) I/O vs the FP32-build
      oracle <= 2e-5; nobsl equal at every point.
  c2  512x512x40, K=64, 2e5 swath + profile obs, and
  c1  the same geometry, K=100, 6 variables with the tracer:
      the batched form with the Gram on MFMA (BF16 hi/lo splits, the chosen
      form, and FP32; BF16 hi/lo with the Householder reduction in FP32),
      the rest FP64, against the plain FP64 form: analysis-mean max error <= 1e-5 relative to the
      variable's max |mean| (the target bound), members <= 1e-4; the plain
      FP64 form against the FP64 oracle <= 1e-10 on 3000 sampled points,
      nobsl equal there.
  c3  256x256x32, K=32, 5e4 obs: the adjoint of the squared misfit of the
      analysis mean to the synthetic truth against central differences
      (h = 1e-4, FP64) on 22 coordinates (obs values and log variances,
      alpha, rho and the localisation scales), <= 1e-4 relative.
The reference's analysis is common_da.f90:22-124.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from helpers import mean_spread_errors, oracle_tables, per_var_err, sample_vs_oracle, to_oracle
from sbpom_letkf_v2_amd import engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(prd, precision, form):
    eng = engine.Engine(engine.make_params(prd, precision, form))
    try:
        a3, a2 = eng.da_main_device(prd)
        torch.cuda.synchronize()
        nl = eng.nobsl().cpu().numpy().T          # (nij1, nlev)
        st = eng.stats()
    finally:
        eng.close()
    return a3, a2, nl, st


def test_c0_full_fp64_every_point():
    _say("case c0 full size, FP64")
    pr = synth.config("c0", device="cpu", real_kind=8)
    rc, a3o, a2o, nlo = po.da_main(to_oracle(pr), dtype=np.float64)
    assert rc == 0
    prd = pr.to("cuda")
    for form in ("plain", "batched"):
        a3, a2, nl, st = _run(prd, "fp64", form)
        assert np.array_equal(nl, nlo), form
        e3 = per_var_err(synth._to_fortran(a3), a3o, axis_var=3)
        e2 = per_var_err(synth._to_fortran(a2), a2o, axis_var=2)
        _say(f"  {form}: {st['points']} points, {st['active_points']} with obs, err3={max(e3):.2e} "
             f"err2={max(e2):.2e}")
        assert st["points"] == 64 * 64 * 16
        assert max(e3) <= 1e-10 and max(e2) <= 1e-10


def test_c0_full_real4_every_point():
    _say("case c0 full size, REAL(4) I/O vs the FP32-build oracle")
    pr = synth.config("c0", device="cpu", real_kind=4)
    rc, a3o, a2o, nlo = po.da_main(to_oracle(pr), dtype=np.float32)
    assert rc == 0
    prd = pr.to("cuda")
    a3, a2, nl, st = _run(prd, "fp64", "plain")
    assert np.array_equal(nl, nlo)
    e3 = per_var_err(synth._to_fortran(a3), a3o, axis_var=3)
    e2 = per_var_err(synth._to_fortran(a2), a2o, axis_var=2)
    _say(f"  err3={max(e3):.2e} err2={max(e2):.2e}")
    assert max(e3) <= 2e-5 and max(e2) <= 2e-5


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", ["c2", "c1"])
def test_full_size_chosen_form_and_fp64_sample(name):
    _say(f"case {name} full size: generating on the GPU")
    prd = synth.config(name, device="cuda", real_kind=8)
    m = prd.meta
    K, nv3, nv2 = m["nmem"], m["nv3d"], m["nv2d"]
    torch.cuda.synchronize()
    _say(f"  {m['nlon']}x{m['nlat']}x{m['nlev']} K={K} nv3d={nv3} nv2d={nv2} nobs={m['nobs']}")
    a3p, a2p, nlp, stp = _run(prd, "fp64", "plain")
    _say(f"  plain FP64: {stp['ms_total']:.0f} ms")
    for prec in ("bf16x3", "fp32", "bf16x3t"):
        a3c, a2c, _, stc = _run(prd, prec, "batched")
        _say(f"  batched {prec}: {stc['ms_total']:.0f} ms, dual points {stc['dual_points']}")
        errs = mean_spread_errors(a3c, a2c, a3p, a2p, nv3, nv2, K)
        for v, e in enumerate(errs):
            _say(f"    var {v}: mean {e['mean']:.2e} spread {e['sprd']:.2e} members {e['member']:.2e}")
        del a3c, a2c
        torch.cuda.empty_cache()
        assert max(e["mean"] for e in errs) <= 1e-5
        assert max(e["member"] for e in errs) <= 1e-4
    # the plain FP64 form against the FP64 oracle on sampled points
    rng = np.random.default_rng(20261204)
    P = m["nlon"] * m["nlat"] * m["nlev"]
    pts = np.concatenate([rng.choice(P, 2500, replace=False),
                          (m["nlev"] - 1) * m["nlon"] * m["nlat"] + rng.choice(m["nlon"] * m["nlat"], 500)])
    pts = np.unique(pts)
    opr = oracle_tables(prd)
    o3, o2, nlo, g3, g2, surf = sample_vs_oracle(prd, opr, (a3p, a2p), pts)
    nij1 = m["nlon"] * m["nlat"]
    assert np.array_equal(nlp[pts % nij1, pts // nij1], nlo)
    e3 = float(np.abs(g3 - o3).max() / np.abs(o3).max())
    e2 = float(np.abs(g2[surf] - o2[surf]).max() / np.abs(o2[surf]).max()) if nv2 else 0.0
    _say(f"  plain FP64 vs oracle on {len(pts)} points ({int((nlo > 0).sum())} with obs, max nobsl "
         f"{int(nlo.max())}): err3={e3:.2e} err2={e2:.2e}")
    assert e3 <= 1e-10 and e2 <= 1e-10
    del a3p, a2p, prd
    torch.cuda.empty_cache()


@pytest.mark.timeout(900)
def test_c3_full_size_adjoint_vs_finite_differences():
    from sbpom_letkf_v2_amd import autograd as ag
    _say("case c3 full size adjoint")
    pr = synth.config("c3", device="cuda", real_kind=8, keep_truth=True)
    m = pr.meta
    t3 = pr.truth3d.reshape(m["nv3d"], m["nlev"], -1)

    def loss(a3, a2):
        return ((a3.mean(1) - t3) ** 2).sum()

    st = ag.AnalysisState(pr, "fp64", "plain")
    try:
        S = lambda x: torch.tensor(float(x), dtype=torch.float64, device="cuda")
        theta = [pr.obsdep.clone(), torch.log(pr.obserr.clone() ** 2), S(pr.params["alpha_rtpp"]),
                 S(pr.params.get("cov_infl_mul", 1.0)), S(pr.params["sigma_obs"]), S(pr.params["sigma_obsv"])]
        th = [t.clone().requires_grad_(True) for t in theta]
        a3, a2 = ag.analysis(st, *th)
        L = loss(a3, a2)
        L.backward()
        g = [t.grad.detach() for t in th]
        gmax = max(float(x.abs().max()) for x in g[:4])
        _say(f"  L={float(L):.6e} |g| dep {float(g[0].abs().max()):.3e} logvar {float(g[1].abs().max()):.3e} "
             f"alpha {float(g[2]):.3e} rho {float(g[3]):.3e} sigma_h {float(g[4]):.3e} sigma_v {float(g[5]):.3e}")

        def value(t):
            with torch.no_grad():
                x3, x2 = ag.analysis(st, *t)
                return float(loss(x3, x2))

        rng = np.random.default_rng(3)
        # the obs with the largest |gradient| and random ones, then alpha and rho
        big = [int(j) for j in torch.topk(g[0].abs(), 4).indices.cpu()]
        picks = [(0, j) for j in big + [int(j) for j in rng.choice(m["nobs"], 5, False)]]
        big = [int(j) for j in torch.topk(g[1].abs(), 4).indices.cpu()]
        picks += [(1, j) for j in big + [int(j) for j in rng.choice(m["nobs"], 5, False)]]
        picks += [(2, None), (3, None), (4, None), (5, None)]
        worst = 0.0
        for which, j in picks:
            # the localisation scales (metres): a step of 1e-5 of their size,
            # the gradient compared with itself (its own units)
            h = 1e-5 * float(theta[which]) if which >= 4 else 1e-4
            tp = [t.clone() for t in theta]
            tm = [t.clone() for t in theta]
            if j is None:
                tp[which] += h
                tm[which] -= h
                ad = float(g[which])
            else:
                tp[which][j] += h
                tm[which][j] -= h
                ad = float(g[which][j])
            fd = (value(tp) - value(tm)) / (2 * h)
            err = abs(fd - ad) / (max(abs(ad), 1e-12) if which >= 4 else max(abs(ad), 1e-3 * gmax))
            worst = max(worst, err)
            _say(f"  d/d{('dep', 'logvar', 'alpha', 'rho', 'sigma_h', 'sigma_v')[which]}[{j}]: adjoint {ad:.6e} "
                 f"fd {fd:.6e} rel {err:.1e}")
        assert worst <= 1e-4
    finally:
        st.close()
        torch.cuda.empty_cache()
