"""Differentiable LETKF analysis on PyTorch-ROCm.

`analysis(state, dep, logvar, alpha_rtpp, cov_infl_mul)` returns the final
analysis ensemble (anal3d, anal2d) of `da_main` (letkf/run/src/common_da.f90:
22-124) as a function of
  dep       the innovations y - H(xf mean) per observation (obsdep),
  logvar    the per-observation log error variance log(obserr^2),
  alpha     the RTPP relaxation factor (common_da.f90:438-461),
  rho       the multiplicative inflation cov_infl_mul (common_letkf.f90:61-210),
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

    def _set_scalars(self, alpha, rho):
        prm = copy.copy(self.params)
        prm.alpha_rtpp = float(alpha)
        prm.cov_infl_mul = float(rho)
        self.eng.set_params(prm)

    def _problem(self, dep, logvar, alpha, rho):
        self._set_scalars(alpha, rho)
        return dataclasses.replace(self.pr, obsdep=dep.detach().contiguous(),
                                   obserr=torch.exp(0.5 * logvar.detach()).contiguous())


class _Analysis(torch.autograd.Function):
    @staticmethod
    def forward(ctx, dep, logvar, alpha, rho, state):
        pr = state._problem(dep, logvar, alpha, rho)
        a3, a2 = state.eng.da_main_device(pr)
        ctx.state = state
        ctx.pr = pr
        ctx.scalars = (float(alpha), float(rho))
        ctx.alpha_like = alpha
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
        return gd, gl, gs[0].to(a.dtype).reshape(a.shape), gs[1].to(a.dtype).reshape(a.shape), None


def analysis(state: AnalysisState, dep, logvar, alpha, rho):
    """(anal3d, anal2d) with autograd through the engine's adjoint; alpha and
    rho are 0-dimensional tensors."""
    return _Analysis.apply(dep, logvar, alpha, rho, state)
