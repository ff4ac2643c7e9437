"""Builders for the benchmark workloads, with phase histories synthesised on
the GPU (libsarf's point-scatterer synthesis, FP64 phase).

  C1  4096 x 2048 samples, Nfft 16384, 2048^2 at 0.05 m, 100 targets (+ clutter)
  C2  1024 x 1024, Nfft 4096, 512^2 at 0.05 m, 20 targets + clutter on every
      4th cell at -20 dB; each pulse carries a line-of-sight error
      a t^2 + b t^3 + c t^4 with a, b, c ~ U(-0.3, 0.3) m (seeded); focusing
      uses the nominal trajectory.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import autofocus as af
from .scenes import CONFIGS, Geometry, ImageGrid, Targets, grid_clutter, uniform


@dataclasses.dataclass
class DeviceScene:
    cfg: object
    geom: Geometry          # nominal geometry (what the radar reports)
    true_geom: Geometry     # geometry the data was synthesised with
    grid: ImageGrid
    targets: Targets
    ph: torch.Tensor        # (Np, K) complex64 on the GPU
    ant: torch.Tensor
    r0: torch.Tensor
    f0: torch.Tensor
    true_coef: np.ndarray | None = None
    n_clutter: int = 0


def _dev(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def synth(ctx, geom: Geometry, tg: Targets, dev="cuda:0", out=None) -> torch.Tensor:
    """Phase history of the scatterers `tg`; many scatterers (clutter) go
    through the non-uniform-FFT synthesis, a few are summed exactly."""
    ant, r0, f0 = _dev(geom.ant, dev), _dev(geom.r0, dev), _dev(geom.f0, dev)
    method = "nufft" if tg.pos.shape[0] > 1000 and geom.K <= 8192 else "direct"
    return ctx.synth_phase_history(ant, r0, f0, geom.df, geom.K, _dev(tg.pos, dev),
                                   _dev(tg.amp.astype(np.complex64), dev), out=out, method=method)


@dataclasses.dataclass
class ServingScene:
    geom: Geometry
    grid: ImageGrid
    nfft: int
    ph: torch.Tensor      # (Np, K) complex64, with the azimuth phase error applied
    ant: torch.Tensor
    r0: torch.Tensor
    f0: torch.Tensor
    phase_err: np.ndarray  # (Np,) applied error (rad)

    @property
    def pixel_pulses(self) -> int:
        return self.grid.Nx * self.grid.Ny * self.geom.Np


def uneven_scene_specs(n: int = 64, seed: int = 4321) -> list:
    """(Np, K, side, n_targets, nfft, phase-error coefficients) of n scenes:
    Np ~ U{256..4096}, K ~ U{256..2048}, side ~ U{128..1024}, 10-50 targets,
    coefficients c_0..c_8 ~ N(0, pi^2); Nfft = 4 * nextpow2(K), <= 16384."""
    from .scenes import normal

    u = uniform(seed, 31, 5 * n).reshape(n, 5)
    specs = []
    for i in range(n):
        Np = 256 + int(u[i, 0] * (4096 - 256 + 1))
        K = 256 + int(u[i, 1] * (2048 - 256 + 1))
        side = 128 + int(u[i, 2] * (1024 - 128 + 1))
        ntg = 10 + int(u[i, 3] * 41)
        nfft = min(16384, 4 * (1 << int(np.ceil(np.log2(K)))))
        c = normal(seed + 13 * i, 41, 9) * np.pi
        specs.append((Np, K, side, ntg, nfft, c))
    return specs


def uneven_scenes(ctx, n: int = 64, seed: int = 4321, spacing: float = 0.05,
                  dev: str = "cuda:0") -> list:
    """n independent scenes (uneven_scene_specs) at `spacing`, each with its
    8th-order polynomial azimuth phase error sum_i c_i t^i (t over [-1, 1])
    multiplied onto every pulse."""
    from .scenes import circular_aperture, point_targets

    out = []
    for i, (Np, K, side, ntg, nfft, c) in enumerate(uneven_scene_specs(n, seed)):
        geom = circular_aperture(Np, K)
        grid = ImageGrid.square(side, spacing)
        tg = point_targets(ntg, 0.8 * spacing * (side - 1), seed + 97 * i, stream=5)
        ph = synth(ctx, geom, tg, dev)
        t = np.linspace(-1.0, 1.0, Np)
        err = np.polyval(c[::-1], t)
        ph *= torch.from_numpy(np.exp(1j * err).astype(np.complex64)).to(dev)[:, None]
        out.append(ServingScene(geom, grid, nfft, ph, _dev(geom.ant, dev), _dev(geom.r0, dev),
                                _dev(geom.f0, dev), err))
    torch.cuda.synchronize()
    return out


def smoothed_random_walk(Np: int, sigma: float, seed: int, stream: int = 21,
                         smooth: float = 16.0) -> np.ndarray:
    """(Np, 3) per-pulse position errors: a Gaussian random walk per axis,
    smoothed by a Gaussian of `smooth` pulses, mean removed, scaled to
    standard deviation `sigma` (m) per axis."""
    from .scenes import normal

    k = np.arange(-int(4 * smooth), int(4 * smooth) + 1)
    ker = np.exp(-0.5 * (k / smooth) ** 2)
    ker /= ker.sum()
    out = np.empty((Np, 3))
    for ax in range(3):
        w = np.cumsum(normal(seed, stream + ax, Np))
        w = np.convolve(np.pad(w, len(k) // 2, mode="edge"), ker, mode="valid")
        w -= w.mean()
        out[:, ax] = w * (sigma / w.std())
    return out


def random_walk_scene(ctx, name: str = "C2", sigma: float = 0.05, seed: int | None = None,
                      clutter: bool = True, dev: str = "cuda:0") -> DeviceScene:
    """The grid-search scene with the error replaced by independent per-pulse
    3-D position errors (smoothed random walk, sigma per axis)."""
    cfg = CONFIGS[name]
    seed = cfg.seed if seed is None else seed
    geom = cfg.geometry()
    grid = cfg.grid()
    tg = cfg.targets()
    err = smoothed_random_walk(geom.Np, sigma, seed)
    true_geom = Geometry(ant=geom.ant + err, r0=geom.r0.copy(), f0=geom.f0.copy(), df=geom.df,
                         K=geom.K)
    ph = synth(ctx, true_geom, tg, dev)
    n_cl = 0
    if clutter and cfg.clutter_stride:
        cl = grid_clutter(grid, cfg.clutter_stride, cfg.clutter_db, seed)
        synth(ctx, true_geom, cl, dev, out=ph)
        n_cl = cl.pos.shape[0]
    torch.cuda.synchronize()
    return DeviceScene(cfg, geom, true_geom, grid, tg, ph, _dev(geom.ant, dev), _dev(geom.r0, dev),
                       _dev(geom.f0, dev), true_coef=None, n_clutter=n_cl)


def los_error_scene(ctx, name: str = "C2", seed: int | None = None, clutter: bool = True,
                    dev: str = "cuda:0", coef_range: float = 0.3) -> DeviceScene:
    cfg = CONFIGS[name]
    seed = cfg.seed if seed is None else seed
    geom = cfg.geometry()
    grid = cfg.grid()
    tg = cfg.targets()
    u = uniform(seed, 11, 3)
    true_coef = (2.0 * u - 1.0) * coef_range
    true_geom = af.perturbed(geom, true_coef, x0=grid.x0, y0=grid.y0)
    ph = synth(ctx, true_geom, tg, dev)
    n_cl = 0
    if clutter and cfg.clutter_stride:
        cl = grid_clutter(grid, cfg.clutter_stride, cfg.clutter_db, seed)
        synth(ctx, true_geom, cl, dev, out=ph)
        n_cl = cl.pos.shape[0]
    torch.cuda.synchronize()
    return DeviceScene(cfg, geom, true_geom, grid, tg, ph, _dev(geom.ant, dev), _dev(geom.r0, dev),
                       _dev(geom.f0, dev), true_coef=true_coef, n_clutter=n_cl)
