"""GPU parity of letkf_lpfgm_core_batched (lpfgm_core, letkf/run/src/
common_lpfgm.f90:444-494) against the oracle per problem, on the same draws.
LPF (mode 2) is bit-exact: the particle picks are comparisons of the same
draws with the same FP64 cumulative weights and the average is a table of
the same running sums.  LPFGM (mode 3) meets 1e-10 of max|W| in FP64 (the
Gram, eigensolve and products run in another order) and 1e-4 in the REAL(4)
kind; problems without rows give W = I."""
import ctypes as C

import numpy as np
import pytest
import torch

import pyoracle as po
from sbpom_letkf_v2_amd import _lib, engine, synth

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _csr(rng, K, sizes, dtype):
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    L = max(int(offs[-1]), 1)
    Y = rng.standard_normal((L, K)) * 0.8
    for b in range(len(sizes)):
        seg = slice(offs[b], offs[b + 1])
        Y[seg] -= Y[seg].mean(1, keepdims=True)
    rd = rng.uniform(0.5, 2.0, L) ** 2
    rl = rng.uniform(0.05, 1.0, L)
    dp = rng.standard_normal(L) * 0.8
    return offs, [np.asarray(a, dtype) for a in (Y, rd, rl, dp)]


def _run(K, sizes, mode, rm, dp, dtype, nmonte=100, gamma=1.3, seed=0):
    rng = np.random.default_rng(seed + K)
    offs, (Y, rd, rl, dep) = _csr(rng, K, sizes, dtype)
    draws = rng.uniform(size=(nmonte, K))
    pr = synth.make_problem(nlon=8, nlat=6, nlev=2, nmem=K, nobs=10, seed=1,
                            real_kind=8 if dtype == np.float64 else 4)
    eng = engine.Engine(engine.make_params(pr))
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a)).cuda()
    tO, tY, trd, trl, tdp = dev(offs), dev(Y), dev(rd), dev(rl), dev(dep)
    nb = len(sizes)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    w = torch.full((nb, K), 7.0, dtype=tdt, device="cuda")
    W = torch.empty((nb, K, K), dtype=tdt, device="cuda")
    vp = lambda t: C.c_void_p(t.data_ptr())
    r = np.ascontiguousarray(draws)
    try:
        _lib.check(_lib.lib().letkf_lpfgm_core_batched(
            eng.h, mode, gamma, nmonte, 1 if rm == "SR" else 0, 1 if dp else 0,
            r.ctypes.data_as(C.POINTER(C.c_double)), nb, vp(tO), vp(tY), vp(trd), vp(trl), vp(tdp), vp(w), vp(W)))
        torch.cuda.synchronize()
    finally:
        eng.close()
    w, W = w.cpu().numpy(), W.cpu().numpy()
    assert not w.any()
    out = []
    for b, n in enumerate(sizes):
        seg = slice(offs[b], offs[b + 1])
        if n == 0:
            np.testing.assert_array_equal(W[b], np.eye(K, dtype=dtype))
            continue
        rc, _, Wo = po.lpfgm_core(Y[seg], rd[seg], rl[seg], dep[seg], draws, mode=mode, cov_infl_gm=gamma, rm=rm,
                                  dp=dp, dtype=dtype)
        assert rc == 0
        out.append((W[b], Wo))
    return out


@pytest.mark.parametrize("K", [20, 64, 100, 128])
@pytest.mark.parametrize("rm,dp", [("MR", True), ("SR", True), ("MR", False)])
def test_lpf_bit_exact(K, rm, dp):
    _say(f"case lpf K={K} {rm} dp={dp}")
    for got, want in _run(K, [0, 1, 3, K - 1, K, 2 * K + 5, 17, 0, 250], 2, rm, dp, np.float64):
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("K", [20, 64, 100, 128])
def test_lpfgm_vs_oracle(K):
    _say(f"case lpfgm K={K}")
    for got, want in _run(K, [0, 1, 3, K - 1, K, 2 * K + 5, 17, 250], 3, "MR", True, np.float64):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_lpfgm_real4():
    _say("case lpf/lpfgm REAL(4) K=40")
    for got, want in _run(40, [0, 5, 39, 40, 90], 2, "MR", True, np.float32):
        np.testing.assert_array_equal(got, want)
    for got, want in _run(40, [0, 5, 39, 40, 90], 3, "MR", True, np.float32):
        assert np.abs(got - want).max() <= 1e-4 * np.abs(want).max()
