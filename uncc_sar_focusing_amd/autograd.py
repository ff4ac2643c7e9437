"""Differentiable image entropy of a backprojected image, for gradient
autofocus.  Forward: libsarf backprojection + entropy; backward: the
hand-written adjoint backprojection (sarf_bp_entropy_grad), which recomputes
the interpolated profiles instead of storing per pixel-pulse values.

PyTorch is used only as the autograd/optimiser harness: every FLOP of the
forward and backward passes runs in libsarf's HIP kernels.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .scenes import ImageGrid


@dataclasses.dataclass
class BpOperator:
    """Fixed inputs of a backprojection whose antenna positions vary."""

    ctx: object
    prof: torch.Tensor   # (Np, Nfft) complex64
    r0: torch.Tensor
    f0: torch.Tensor
    grid: ImageGrid
    maxWr: float
    los: tuple[float, float] | None = None

    def image(self, ant: torch.Tensor) -> torch.Tensor:
        return self.ctx.backproject(self.prof, ant.detach().contiguous(), self.r0, self.f0,
                                    self.grid, self.maxWr, los=self.los)


class BackprojectEntropy(torch.autograd.Function):
    """E(ant) = entropy(BP(prof; ant)); d E / d ant from the adjoint kernel."""

    @staticmethod
    def forward(fctx, ant: torch.Tensor, op: BpOperator):
        ant_c = ant.detach().contiguous()
        img = op.image(ant_c)
        ent = op.ctx.image_entropy(img)
        fctx.save_for_backward(ant_c, img, ent)
        fctx.op = op
        return ent[0].clone()

    @staticmethod
    def backward(fctx, g_e):
        ant_c, img, ent = fctx.saved_tensors
        op = fctx.op
        grad = op.ctx.entropy_grad(op.prof, ant_c, op.r0, op.f0, op.grid, op.maxWr, img, ent,
                                   los=op.los)
        return grad * g_e, None


def entropy(ant: torch.Tensor, op: BpOperator) -> torch.Tensor:
    return BackprojectEntropy.apply(ant, op)


def smoothing_operator(Np: int, sigma: float, device) -> "callable":
    """p = S q: a Gaussian smoothing of `sigma` pulses along slow time (rows
    of S normalised, so constants pass unchanged), as a differentiable map of
    (Np, k) parameter tensors.  Descending over q instead of the per-pulse
    offsets p keeps every update inside the band of a smooth trajectory
    error; the unknowns stay one per pulse and axis."""
    i = torch.arange(Np, dtype=torch.float64, device=device)
    S = torch.exp(-0.5 * ((i[:, None] - i[None, :]) / sigma) ** 2)
    S = S / S.sum(dim=1, keepdim=True)
    return lambda q: S @ q


def los_residual_rms(err: np.ndarray, direction: np.ndarray) -> float:
    """rms of the line-of-sight component err_n . dir_n of per-pulse
    position errors (Np, 3) after removing its least-squares line over
    normalised slow time: a constant or linear LOS offset only shifts the
    image, so it is not observable by a focus metric."""
    r = np.sum(np.asarray(err) * np.asarray(direction), axis=1)
    t = np.linspace(-1.0, 1.0, r.size)
    A = np.stack([np.ones_like(t), t], axis=1)
    c, *_ = np.linalg.lstsq(A, r, rcond=None)
    return float(np.sqrt(np.mean((r - A @ c) ** 2)))


@dataclasses.dataclass
class GradientResult:
    params: torch.Tensor
    entropies: np.ndarray   # E before each step
    gpu_ms: np.ndarray      # cumulative device time at the end of each step
    iters: int


def gradient_autofocus(op: BpOperator, make_ant, params0: torch.Tensor, iters: int,
                       lr: float, betas=(0.9, 0.999)) -> GradientResult:
    """Adam descent of E(make_ant(params)).  No host synchronisation inside the
    loop: entropies and per-step device timestamps are read at the end."""
    p = params0.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([p], lr=lr, betas=betas)
    ents = torch.empty(iters, dtype=torch.float64, device=p.device)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(iters + 1)]
    evs[0].record()
    for i in range(iters):
        opt.zero_grad(set_to_none=True)
        e = entropy(make_ant(p), op)
        e.backward()
        opt.step()
        ents[i] = e.detach()
        evs[i + 1].record()
    evs[-1].synchronize()
    ms = np.array([evs[0].elapsed_time(evs[i + 1]) for i in range(iters)])
    return GradientResult(params=p.detach(), entropies=ents.cpu().numpy(), gpu_ms=ms, iters=iters)
