"""Differentiable LETKF analysis on PyTorch-ROCm.

`analysis(state, dep, logvar, alpha_rtpp, cov_infl_mul)` returns the final
analysis ensemble (anal3d, anal2d) of `da_main` (letkf/run/src/common_da.f90:
22-124) as a function of
  dep       the innovations y - H(xf mean) per observation (obsdep),
  logvar    the per-observation log error variance log(obserr^2),
  alpha     the RTPP relaxation factor (common_da.f90:438-461),
  rho       the multiplicative inflation cov_infl_mul (common_letkf.f90:61-210),
  sigma_h, sigma_v (optional) the localisation scales sigma_obs, sigma_obsv
            [m] of rloc (common_da.f90:282-287), the cut-offs dist_zero(v)
            of the state held fixed,
with the grid, the forecast ensemble and H(dXf) held fixed in `state`.  The
forward pass is the engine's analysis (letkf_da_main_device); the backward
pass is the engine's hand-written adjoint (letkf_da_adjoint_device), which
recomputes every point's eigen-decomposition in FP64 instead of storing it.
Gradients flow to every tensor argument that requires them.
"""
from __future__ import annotations

import copy
import dataclasses

import torch

from . import engine


class AnalysisState:
    """A problem (device tensors, real_kind 8) and one engine handle."""

    def __init__(self, pr, precision="fp64", form="plain"):
        if pr.fcst3d.dtype != torch.float64 or not pr.fcst3d.is_cuda:
            raise TypeError("the differentiable analysis needs FP64 tensors on the GPU")
        self.pr = pr
        self.params = engine.make_params(pr, precision, form)
        self.eng = engine.Engine(self.params)

    def close(self):
        self.eng.close()

    def _set_scalars(self, alpha, rho, sigma_h=None, sigma_v=None):
        prm = copy.copy(self.params)
        prm.alpha_rtpp = float(alpha)
        prm.cov_infl_mul = float(rho)
        if sigma_h is not None:
            prm.sigma_obs = float(sigma_h)
        if sigma_v is not None:
            prm.sigma_obsv = float(sigma_v)
        self.eng.set_params(prm)

    def _problem(self, dep, logvar, alpha, rho, sigma_h=None, sigma_v=None):
        self._set_scalars(alpha, rho, sigma_h, sigma_v)
        return dataclasses.replace(self.pr, obsdep=dep.detach().contiguous(),
                                   obserr=torch.exp(0.5 * logvar.detach()).contiguous())


class _Analysis(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dep, logvar, alpha, rho, sigma_h, sigma_v, state):
        pr = state._problem(dep, logvar, alpha, rho, sigma_h, sigma_v)
        a3, a2 = state.eng.da_main_device(pr)
        ctx.state = state
        ctx.pr = pr
        ctx.scalars = (float(alpha), float(rho), None if sigma_h is None else float(sigma_h),
                       None if sigma_v is None else float(sigma_v))
        ctx.alpha_like = alpha
        ctx.has_sigma = (sigma_h is not None, sigma_v is not None)
        return a3, a2

    @staticmethod
    def backward(ctx, g3, g2):
        state = ctx.state
        state._set_scalars(*ctx.scalars)
        if g3 is None:
            g3 = torch.zeros_like(ctx.pr.fcst3d)
        if g2 is None:
            g2 = torch.zeros_like(ctx.pr.fcst2d)
        gd, gl, gs = state.eng.da_adjoint_device(ctx.pr, g3, g2)
        a = ctx.alpha_like
        sc = lambda x: x.to(a.dtype).reshape(a.shape)
        gh = sc(gs[2]) if ctx.has_sigma[0] else None
        gv = sc(gs[3]) if ctx.has_sigma[1] else None
        return gd, gl, sc(gs[0]), sc(gs[1]), gh, gv, None


def analysis(state: AnalysisState, dep, logvar, alpha, rho, sigma_h=None, sigma_v=None):
    """(anal3d, anal2d) with autograd through the engine's adjoint; alpha,
    rho and the optional localisation scales sigma_h, sigma_v [m] are
    0-dimensional tensors (without them the state's scales are used and
    receive no gradient)."""
    return _Analysis.apply(dep, logvar, alpha, rho, sigma_h, sigma_v, state)
