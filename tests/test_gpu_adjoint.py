"""Adjoint of the analysis (letkf_da_adjoint_device / autograd.analysis):
central finite differences in FP64 and an independent FP64 reference
gradient (a dense torch restatement of letkf_core + RTPP over the oracle's
local observation sets, differentiated by torch.autograd through
torch.linalg.eigh on the CPU), with respect to the innovations, the per-obs
log error variances, RTPP alpha, rho and the localisation scales
sigma_obs / sigma_obsv (cut-offs held fixed); and run-to-run bit equality
of the gradients (no atomics in the adjoint)."""
import numpy as np
import pytest
import torch

from helpers import to_oracle
from sbpom_letkf_v2_amd import autograd as ag
from sbpom_letkf_v2_amd import synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _problem(K=16, nobs=400, seed=11, nv2d=1, alpha=0.7):
    return synth.make_problem(nlon=24, nlat=20, nlev=5, nmem=K, nv3d=4, nv2d=nv2d, nobs=nobs, seed=seed,
                              alpha_rtpp=alpha, sigma_h_cells=1.5, sigma_v=1.0, real_kind=8,
                              keep_truth=True, device="cuda")


def _params(pr):
    dep = pr.obsdep.clone()
    logvar = torch.log(pr.obserr.clone() ** 2)
    S = lambda x: torch.tensor(float(x), dtype=torch.float64, device="cuda")
    return (dep, logvar, S(pr.params["alpha_rtpp"]), S(pr.params.get("cov_infl_mul", 1.0)),
            S(pr.params["sigma_obs"]), S(pr.params["sigma_obsv"]))


NAMES = ("dep", "logvar", "alpha", "rho", "sigma_h", "sigma_v")


def _linear_loss(R3, R2):
    def f(a3, a2):
        s = (R3 * a3).sum()
        return s + (R2 * a2).sum() if R2 is not None else s
    return f


def _mean_misfit(pr):
    """Squared misfit of the analysis mean to the synthetic truth (the loss of
    the differentiable workload)."""
    t3 = pr.truth3d.reshape(pr.truth3d.shape[0], pr.truth3d.shape[1], -1)

    def f(a3, a2):
        return ((a3.mean(1) - t3) ** 2).sum()
    return f


def _grads(st, theta, loss):
    th = [t.clone().requires_grad_(True) for t in theta]
    a3, a2 = ag.analysis(st, *th)
    L = loss(a3, a2)
    L.backward()
    return L.item(), [t.grad.clone() for t in th]


def _value(st, theta, loss):
    with torch.no_grad():
        a3, a2 = ag.analysis(st, *theta)
        return float(loss(a3, a2))


def _fd_check(st, theta, loss, g, ncoord, seed, h=1e-4):
    """Worst |fd - adjoint| / max(|adjoint|, 1e-3 max|gradient|) over random
    coordinates of each argument (the floor keeps a gradient that is zero by
    construction, such as d(mean misfit)/d alpha, from dividing by zero)."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    gmax = max(float(x.abs().max()) for x in g[:4])
    for which in range(len(theta)):
        n = theta[which].numel()
        coords = [()] if theta[which].dim() == 0 else [(int(j),) for j in rng.choice(n, min(ncoord, n), False)]
        # the localisation scales are metres: a step of 1e-5 of their size,
        # their gradient compared with itself (it has its own units)
        loc = which >= 4
        hw = 1e-5 * float(theta[which]) if loc else h
        for c in coords:
            tp = [t.clone() for t in theta]
            tm = [t.clone() for t in theta]
            tp[which][c] += hw
            tm[which][c] -= hw
            fd = (_value(st, tp, loss) - _value(st, tm, loss)) / (2 * hw)
            ad = float(g[which][c])
            scale = max(abs(ad), 1e-12) if loc else max(abs(ad), 1e-3 * gmax, 1e-12)
            err = abs(fd - ad) / scale
            if which >= 2:
                _say(f"    d/d{NAMES[which]}: adjoint {ad:.6e} fd {fd:.6e} rel {err:.1e}")
            worst = max(worst, err)
    return worst


@pytest.mark.parametrize("K,nv2d", [(16, 1), (32, 0)])
def test_adjoint_matches_finite_differences(K, nv2d):
    _say(f"case adjoint FD K={K} nv2d={nv2d}")
    pr = _problem(K=K, nv2d=nv2d)
    st = ag.AnalysisState(pr)
    try:
        theta = _params(pr)
        gen = torch.Generator(device="cuda").manual_seed(5)
        R3 = torch.randn(pr.fcst3d.shape, generator=gen, device="cuda", dtype=torch.float64)
        R2 = torch.randn(pr.fcst2d.shape, generator=gen, device="cuda", dtype=torch.float64) if nv2d else None
        for name, loss in (("linear", _linear_loss(R3, R2)), ("mean misfit", _mean_misfit(pr))):
            L, g = _grads(st, theta, loss)
            worst = _fd_check(st, theta, loss, g, ncoord=12, seed=K)
            _say(f"  {name}: L={L:.6e} |g| dep {float(g[0].abs().max()):.3e} logvar "
                 f"{float(g[1].abs().max()):.3e} alpha {float(g[2]):.3e} rho {float(g[3]):.3e} "
                 f"sigma_h {float(g[4]):.3e} sigma_v {float(g[5]):.3e}; worst FD rel err {worst:.2e}")
            assert worst <= 1e-4
    finally:
        st.close()


def _reference_loss(pn, opr, local, dep, logvar, alpha, rho, sh, sv, R3, R2):
    """Dense FP64 restatement: letkf_core (common_letkf.f90:61-210) with
    torch.linalg.eigh, Analysis_Ensemble and RTPP (common_da.f90:407-461);
    rloc = exp(-((dist/sh)^2 + (dlev/sv)^2)/2) (common_da.f90:282-287) on
    the oracle's local sets, with dist^2 and dlev^2 of each pair recovered
    from the oracle's rloc at the problem's own scales and the vertical
    distance rule of the search (SSH: |depth| below the surface, 0 at it)."""
    import pyoracle as po  # noqa: F401  (the local sets come from the oracle search)
    nij1, nlev, K, nv3 = pn.fcst3d.shape
    nv2 = pn.fcst2d.shape[2] if pn.fcst2d.ndim == 3 else 0
    X3 = torch.as_tensor(np.ascontiguousarray(pn.fcst3d))
    X2 = torch.as_tensor(np.ascontiguousarray(pn.fcst2d)) if nv2 else None
    Yall = torch.as_tensor(np.ascontiguousarray(pn.obshdxf))
    eye = torch.eye(K, dtype=torch.float64)
    L = torch.zeros((), dtype=torch.float64)
    for k in range(nlev):
        for i in range(nij1):
            X = X3[i, k]
            R = R3[i, k]
            if k == nlev - 1 and nv2:
                X = torch.cat([X, X2[i]], 1)
                R = torch.cat([R, R2[i]], 1)
            xm = X.mean(0)
            dX = X - xm
            idx, rl = local[(i, k)]
            if len(idx):
                ix = torch.as_tensor(idx.astype(np.int64))
                Y = Yall[ix]
                sh0, sv0 = float(pn.params["sigma_obs"]), float(pn.params["sigma_obsv"])
                rdep = float(pn.depth1[i, k])
                el = pn.obselm[idx]
                dlev = np.where(el == pn.params["id_z_obs"], abs(rdep) if k < nlev - 1 else 0.0,
                                np.abs(pn.obslev[idx] - rdep))
                dlev2 = torch.as_tensor(dlev ** 2)
                dist2 = torch.as_tensor((-2.0 * np.log(rl) - dlev ** 2 / sv0 ** 2) * sh0 ** 2)
                rloc = torch.exp(-0.5 * (dist2 / sh ** 2 + dlev2 / sv ** 2))
                q = rloc / torch.exp(logvar[ix])
                C = (K - 1) / rho * eye + Y.T @ (q[:, None] * Y)
                b = Y.T @ (q * dep[ix])
                lam, V = torch.linalg.eigh(C)
                w = V @ ((V.T @ b) / lam)
                W = (K - 1) ** 0.5 * (V * lam.rsqrt()) @ V.T
                xa = xm + w @ dX + W @ dX
            else:
                xa = xm + torch.sqrt(rho) * dX
            am = xa.mean(0)
            out = am + alpha * dX + (1 - alpha) * (xa - am)
            L = L + (R * out).sum()
    return L


def test_adjoint_matches_fp64_reference_gradient():
    import pyoracle as po
    _say("case adjoint vs torch.linalg.eigh reference")
    pr = synth.make_problem(nlon=14, nlat=12, nlev=3, nmem=12, nv3d=4, nv2d=1, nobs=160, seed=21,
                            alpha_rtpp=0.6, cov_infl_mul=1.1, sigma_h_cells=1.5, sigma_v=1.0,
                            real_kind=8, device="cuda")
    pn = pr.numpy()
    opr = to_oracle(pn)
    nij1, nlev = pn.depth1.shape
    local = {}
    for k in range(nlev):
        for i in range(nij1):
            idx, _, _, rl, _ = po.obs_local(opr, i, k)
            local[(i, k)] = (idx, rl)
    st = ag.AnalysisState(pr)
    try:
        theta = _params(pr)
        gen = torch.Generator(device="cuda").manual_seed(9)
        R3t = torch.randn(pr.fcst3d.shape, generator=gen, device="cuda", dtype=torch.float64)
        R2t = torch.randn(pr.fcst2d.shape, generator=gen, device="cuda", dtype=torch.float64)
        L, g = _grads(st, theta, _linear_loss(R3t, R2t))
        # the same loss on the CPU: tensors in Fortran index order [i, k, m, v] / [i, m, v]
        R3 = R3t.permute(3, 2, 1, 0).cpu()
        R2 = R2t.permute(2, 1, 0).cpu()
        th = [t.detach().cpu().clone().requires_grad_(True) for t in theta]
        Lr = _reference_loss(pn, opr, local, *th, R3, R2)
        Lr.backward()
        _say(f"  L gpu {L:.12e} ref {float(Lr):.12e}")
        assert abs(L - float(Lr)) <= 1e-10 * max(1.0, abs(float(Lr)))
        for name, ga, gr in zip(NAMES, g, [t.grad for t in th]):
            err = float((ga.cpu() - gr).abs().max()) / max(float(gr.abs().max()), 1e-300)
            _say(f"  d/d{name}: max rel err {err:.2e}")
            assert err <= 1e-8
    finally:
        st.close()


def test_adjoint_is_bit_reproducible():
    """Two adjoint passes of the same analysis give the same bits (the
    per-obs sums run in point order after a stable sort, the scalar sums in
    a fixed block order)."""
    _say("case adjoint determinism")
    pr = _problem(K=32, nobs=900, seed=13)
    st = ag.AnalysisState(pr)
    try:
        theta = _params(pr)
        gen = torch.Generator(device="cuda").manual_seed(3)
        R3 = torch.randn(pr.fcst3d.shape, generator=gen, device="cuda", dtype=torch.float64)
        R2 = torch.randn(pr.fcst2d.shape, generator=gen, device="cuda", dtype=torch.float64)
        _, g1 = _grads(st, theta, _linear_loss(R3, R2))
        _, g2 = _grads(st, theta, _linear_loss(R3, R2))
        for name, a, b in zip(NAMES, g1, g2):
            assert torch.equal(a, b), name
    finally:
        st.close()
