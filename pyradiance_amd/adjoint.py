"""Differentiable sensor irradiance: rcontrib on the GPU plus its path-replay
adjoint, behind a torch.autograd.Function.

The forward pass is rcontrib with binned sky modifiers (the coefficient
matrix DC[sensor][bin][channel], rt/rcontrib.c:272-338) followed by the
product with the bin radiances L[bin][channel]:

    E[s][c] = sum_b DC[s][b][c] * L[b][c]

Gradients:
  * with respect to L: DC itself (E is linear in L), so dLoss/dL[b][c] =
    sum_s dLoss/dE[s][c] * DC[s][b][c];
  * with respect to the plastic arguments of chosen material groups (red,
    green, blue reflectance, specularity, roughness): rg_contrib_adjoint, a
    second pass over the same rays with the same random numbers that adds,
    at every binned deposit, the derivative of the deposit's path
    coefficient weighted by dLoss/dE and L.  Per-bounce state is
    recomputed by the replay, not stored.

PyTorch is used here only for autograd; the work runs in libradgpu.so.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib


class DiffScene:
    """A committed scene, its sensors and tracked sky modifiers, and the
    material groups whose arguments are differentiated."""

    def __init__(self, rad: str, sensors: np.ndarray, groups: list, mods: list, params,
                 flags: int = _lib.RT_IMM_IRRAD, device: int = 0):
        self.ctx = _lib.Context(device)
        self.ctx.load_text(rad).commit()
        self.ctx.set_params(params)
        self.rays = np.ascontiguousarray(sensors, dtype=np.float64).reshape(-1, 6)
        self.mods = mods
        self.flags = flags
        self.groups = list(groups)
        self.mats = [self.ctx.material_index(g) for g in self.groups]
        for m in self.mats:
            kind, _ = self.ctx.material_get(m)
            if kind != 0:
                raise ValueError("only plastic material groups can be differentiated")
        self.nbins = sum(m.nbins for m in mods)

    def set_theta(self, theta: np.ndarray) -> None:
        """Plastic arguments per group, (ngroups, 5); negative values clip to 0."""
        theta = np.asarray(theta, dtype=np.float64).reshape(len(self.mats), 5)
        for m, th in zip(self.mats, np.maximum(theta, 0.0)):
            self.ctx.material_set(m, th)

    def coefficients(self) -> np.ndarray:
        """DC matrix (nsensors, nbins, 3) for the current arguments."""
        dc = self.ctx.contrib(self.rays, self.mods, 1, self.flags)
        return dc.reshape(self.rays.shape[0], self.nbins, 3)

    def adjoint(self, dlde: np.ndarray, sky: np.ndarray) -> np.ndarray:
        """dLoss/dtheta (ngroups, 5) for dLoss/dE = dlde (nsensors, 3)."""
        return self.ctx.contrib_adjoint(self.rays, self.mods, dlde, np.asarray(sky, np.float64).ravel(),
                                        self.mats, 1, self.flags)

    def close(self):
        self.ctx.close()


class Irradiance(torch.autograd.Function):
    """E = irradiance(theta, sky): theta (ngroups, 5), sky (nbins, 3) -> (nsensors, 3)."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, sky: torch.Tensor, scene: DiffScene):
        scene.set_theta(theta.detach().cpu().numpy())
        dc = torch.from_numpy(scene.coefficients()).to(device=sky.device, dtype=torch.float32)
        e = (dc * sky.detach().to(torch.float32)[None]).sum(1)
        ctx.scene = scene
        ctx.theta_dtype = theta.dtype
        ctx.save_for_backward(dc, sky)
        return e.to(sky.dtype)

    @staticmethod
    def backward(ctx, ge: torch.Tensor):
        dc, sky = ctx.saved_tensors
        scene = ctx.scene
        gsky = gtheta = None
        if ctx.needs_input_grad[1]:
            gsky = (dc * ge.to(torch.float32)[:, None, :]).sum(0).to(sky.dtype)
        if ctx.needs_input_grad[0]:
            g = scene.adjoint(ge.detach().to(torch.float64).cpu().numpy(), sky.detach().cpu().numpy())
            gtheta = torch.from_numpy(g).to(device=ge.device, dtype=ctx.theta_dtype)
        return gtheta, gsky, None


def irradiance(theta: torch.Tensor, sky: torch.Tensor, scene: DiffScene) -> torch.Tensor:
    return Irradiance.apply(theta, sky, scene)
