"""Pins the oracle's lpfgm_core (letkf/run/src/common_lpfgm.f90:444-494).

common_lpfgm.f90 uses common_mpi (MPI) and is not buildable here, so the
oracle is checked against what the algorithm fixes:
  * particle weights: resampling columns placed on either side of numpy's
    softmax cumulative weights of -0.5 sum (d - y_j)^2 rloc / rdiag land in
    the predicted particles (LPF, one resampling, no diagonal priority);
  * equal weights with one draw per weight bin and diagonal priority give
    the identity; one dominant particle takes every column;
  * systematic resampling reproduces the reference's in-place update of
    rand(1) (:115-117), which differs from the textbook offsets;
  * every resampling assigns each column exactly once (column sums of the
    LPF matrix = 1 to rounding);
  * the Gaussian-mixture factor equals the closed form alpha Pa + (Pa b) 1^T
    (alpha = (m-1)/gamma, Pa = (Y^T R^-1 Y + alpha I)^-1, b = Y^T R^-1 d),
    and LPFGM = W_GM W_LPF.
"""
import numpy as np
import pytest

import pyoracle as po


def _problem(rng, n, K, spread=1.0):
    hdxf = rng.standard_normal((n, K)) * spread
    hdxf -= hdxf.mean(axis=1, keepdims=True)
    rdiag = rng.uniform(0.5, 2.0, n) ** 2
    rloc = rng.uniform(0.2, 1.0, n)
    dep = rng.standard_normal(n) * 0.7
    return hdxf, rdiag, rloc, dep


def _cw(hdxf, rdiag, rloc, dep):
    ll = -0.5 * (((dep[:, None] - hdxf) ** 2) * (rloc / rdiag)[:, None]).sum(axis=0)
    w = np.exp(ll - ll.max())
    return np.cumsum(w / w.sum())


def test_weights_through_resampling_columns():
    rng = np.random.default_rng(81)
    n, K = 12, 16
    hdxf, rdiag, rloc, dep = _problem(rng, n, K, spread=0.6)
    cw = _cw(hdxf, rdiag, rloc, dep)
    # column j's draw sits 1e-9 below or above a cumulative weight: the
    # particle picked is the first i with draw <= cw_i
    targets = rng.integers(0, K - 1, K)
    above = rng.uniform(size=K) < 0.5
    draws = np.where(above, cw[targets] + 1e-9, cw[targets] - 1e-9)
    want_row = np.where(above, targets + 1, targets)
    # multinomial resampling sorts the draws; sort here so column j keeps its draw
    order = np.argsort(draws)
    draws, want_row = draws[order], want_row[order]
    rc, w, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, draws[None, :], mode=2, dp=False)
    assert rc == 0 and np.all(w == 0)
    got_row = W.argmax(axis=0)
    np.testing.assert_array_equal(got_row, want_row)
    assert np.all(W.max(axis=0) == 1.0) and np.all((W != 0).sum(axis=0) == 1)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_equal_weights_identity_and_dominant_particle(dtype):
    rng = np.random.default_rng(82)
    n, K, nm = 5, 10, 7
    hdxf = np.zeros((n, K))
    rdiag, rloc, dep = np.ones(n), np.ones(n), rng.standard_normal(n)
    mid = (np.arange(K) + 0.5) / K                       # one draw in each weight bin
    draws = np.tile(mid, (nm, 1))
    for m in range(nm):
        rng.shuffle(draws[m])                              # sorted by the core (MR)
    rc, _, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, draws, mode=2, dp=True, dtype=dtype)
    acc = dtype(0)
    for _ in range(nm):                                    # the running average of 1/nmonte
        acc = dtype(np.float64(acc) + 1.0 / nm)
    np.testing.assert_array_equal(W, np.eye(K, dtype=dtype) * acc)
    # member 3 matches the innovations exactly, the others are far off
    hdxf = np.full((n, K), 40.0)
    hdxf[:, 3] = dep
    rc, _, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, rng.uniform(size=(nm, K)), mode=2, dp=True, dtype=dtype)
    want = np.zeros((K, K), dtype)
    want[3, :] = acc
    np.testing.assert_array_equal(W, want)


def test_systematic_resampling_in_place_offsets():
    # weights 0.1, 0.25, 0.25, 0.4 from one obs: w_j ~ exp(-y_j^2 / 2)
    target = np.array([0.1, 0.25, 0.25, 0.4])
    y = np.sqrt(-2.0 * np.log(target / target.max()))
    hdxf, rdiag, rloc, dep = y[None, :], np.ones(1), np.ones(1), np.zeros(1)
    u = 0.99
    # rand(1) <- u/4 first, then rand(i) = (rand(1) + i - 1)/4 with the new rand(1):
    # draws 0.2475, 0.311875, 0.561875, 0.811875 against cw 0.1, 0.35, 0.6, 1.0
    _, _, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, np.array([[u, 0.5, 0.5, 0.5]]), mode=2, rm="SR", dp=False)
    np.testing.assert_array_equal(W.argmax(axis=0), [1, 1, 2, 3])
    # (the textbook offsets (u + i - 1)/4 would give particles 1, 2, 3, 3)


@pytest.mark.parametrize("dp", [True, False])
def test_columns_assigned_once(dp):
    rng = np.random.default_rng(83)
    n, K, nm = 30, 24, 50
    hdxf, rdiag, rloc, dep = _problem(rng, n, K)
    _, _, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, rng.uniform(size=(nm, K)), mode=2, dp=dp)
    np.testing.assert_allclose(W.sum(axis=0), 1.0, atol=1e-13)
    counts = np.rint(W * nm)
    np.testing.assert_allclose(W * nm, counts, atol=1e-11)
    if dp:   # diagonal priority: a particle drawn anywhere keeps its own column
        drawn = counts.sum(axis=1) > 0
        assert np.all(np.diag(counts)[drawn] > 0)


@pytest.mark.parametrize("gamma", [1.0, 1.7])
def test_gm_closed_form_and_product(gamma):
    rng = np.random.default_rng(84)
    n, K, nm = 40, 20, 30
    hdxf, rdiag, rloc, dep = _problem(rng, n, K)
    draws = rng.uniform(size=(nm, K))
    _, _, Wl = po.lpfgm_core(hdxf, rdiag, rloc, dep, draws, mode=2, cov_infl_gm=gamma)
    rc, w, W = po.lpfgm_core(hdxf, rdiag, rloc, dep, draws, mode=3, cov_infl_gm=gamma)
    assert rc == 0 and np.all(w == 0)
    Rl = rloc / rdiag
    alpha = (K - 1) / gamma
    Pa = np.linalg.inv(hdxf.T @ (Rl[:, None] * hdxf) + alpha * np.eye(K))
    wbar = Pa @ (hdxf.T @ (Rl * dep))
    Wgm = alpha * Pa + np.outer(wbar, np.ones(K))
    np.testing.assert_allclose(W, Wgm @ Wl, rtol=0, atol=1e-11)
    # the FP32 build within single-precision rounding
    _, _, W32 = po.lpfgm_core(hdxf, rdiag, rloc, dep, draws, mode=3, cov_infl_gm=gamma, dtype=np.float32)
    np.testing.assert_allclose(W32, W, rtol=0, atol=5e-5)
