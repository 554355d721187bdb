"""Differentiable sensor irradiance: rcontrib on the GPU plus its path-replay
adjoint, behind a torch.autograd.Function.

The forward pass is rcontrib with binned sky modifiers (the coefficient
matrix DC[sensor][bin][channel], rt/rcontrib.c:272-338) followed by the
product with the bin radiances L[bin][channel]:

    E[s][c] = sum_b DC[s][b][c] * L[b][c]

Gradients:
  * with respect to L: DC itself (E is linear in L), so dLoss/dL[b][c] =
    sum_s dLoss/dE[s][c] * DC[s][b][c];
  * with respect to the real arguments of chosen plastic, metal or trans
    material groups (red, green, blue reflectance, specularity, roughness,
    and for trans transmissivity and transmitted specularity):
    rg_contrib_adjoint, a second pass over the same rays with the same
    random numbers that adds, at every binned deposit (shadow-ray deposits
    included), the derivative of the deposit's path coefficient weighted by
    dLoss/dE and L.  The roughness derivative adds the likelihood-ratio
    score of the Gaussian lobe samples (rt/normal.c:364-495), so it is exact
    in expectation rather than sample by sample.  Per-bounce state is
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
        self.nargs = []
        for m in self.mats:
            kind, _ = self.ctx.material_get(m)
            if kind not in (0, 1, 2):
                raise ValueError("only plastic, metal and trans material groups can be differentiated")
            self.nargs.append(7 if kind == 2 else 5)
        self.width = max(self.nargs, default=5)
        self.nbins = sum(m.nbins for m in mods)

    def set_theta(self, theta: np.ndarray) -> None:
        """Material arguments per group, (ngroups, 5) or (ngroups, 7) when a
        group is trans (plastic and metal groups use the first 5 columns);
        negative values clip to 0."""
        theta = np.asarray(theta, dtype=np.float64).reshape(len(self.mats), self.width)
        for m, k, th in zip(self.mats, self.nargs, np.maximum(theta, 0.0)):
            self.ctx.material_set(m, th[:k])

    def coefficients(self) -> np.ndarray:
        """DC matrix (nsensors, nbins, 3) for the current arguments."""
        dc = self.ctx.contrib(self.rays, self.mods, 1, self.flags)
        return dc.reshape(self.rays.shape[0], self.nbins, 3)

    def adjoint(self, dlde: np.ndarray, sky: np.ndarray) -> np.ndarray:
        """dLoss/dtheta (ngroups, 5 or 7, as set_theta) for dLoss/dE = dlde (nsensors, 3)."""
        g = self.ctx.contrib_adjoint(self.rays, self.mods, dlde, np.asarray(sky, np.float64).ravel(),
                                     self.mats, 1, self.flags)
        return g[:, :self.width]

    def close(self):
        self.ctx.close()


class Irradiance(torch.autograd.Function):
    """E = irradiance(theta, sky): theta (ngroups, 5 or 7), sky (nbins, 3) -> (nsensors, 3)."""

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
