"""Autofocus drivers on top of libsarf: batched grid search over trajectory
candidates (all candidate images in one persistent launch, scored by image
entropy), with the trajectory model of a line-of-sight (LOS) polynomial.

Reference counterparts: the cuGridSearch-driven search of
src/autofocus/autofocus_gpu.cu:707-1260 (candidate trajectory model
gridSearchErrorFunctions_dft.cuh:279-291, multi-resolution shrink
autofocus_gpu.cu:938-939) -- here the model is the LOS polynomial of the
batched workload and the objective the entropy E = -sum p ln p.
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np
import torch

from .scenes import Geometry


def los_directions(ant: np.ndarray, x0: float = 0.0, y0: float = 0.0) -> np.ndarray:
    """(Np, 3) unit vectors from the scene centre (x0, y0, 0) to each phase
    centre -- the line of sight a range error moves the antenna along."""
    u = ant - np.array([x0, y0, 0.0])
    return u / np.linalg.norm(u, axis=1, keepdims=True)


def poly_basis(Np: int, powers=(2, 3, 4)) -> np.ndarray:
    """(len(powers), Np): t^p over normalised slow time t = linspace(-1, 1, Np)."""
    t = np.linspace(-1.0, 1.0, Np)
    return np.stack([t ** p for p in powers])


def apply_los_offsets(ant: np.ndarray, coef, basis: np.ndarray, direction: np.ndarray) -> np.ndarray:
    """Antenna positions a_n + (coef . basis[:, n]) dir_n (FP64)."""
    off = np.asarray(coef, dtype=np.float64) @ basis
    return ant + off[:, None] * direction


def reference_trajectory_candidates(params, D: int, Np: int):
    """The reference's grid-search candidates -- absolute phase centres that
    are polynomials in the pulse index i along x, y and z
    (src/autofocus/gridSearchErrorFunctions_dft.cuh:279-291; D = 7: linear
    with a shared velocity scale p6; D = 10: quadratic with scale p9) -- as
    the linear per-basis form of batch_entropy: with antenna positions 0,
    a_i = sum_k coef_k basis_k[i] dir_k.  params: (ncand, D).  Returns
    (basis (nb, Np), direction (nb, Np, 3), coefs (ncand, nb)); nb = 6 for
    D = 7 (1, i per axis), 9 for D = 10 (1, i, i^2 per axis)."""
    p = np.atleast_2d(np.asarray(params, dtype=np.float64))
    if p.shape[1] != D or D not in (7, 10):
        raise ValueError("params must be (ncand, 7) or (ncand, 10)")
    i = np.arange(Np, dtype=np.float64)
    powers = (0, 1) if D == 7 else (0, 1, 2)
    nb = 3 * len(powers)
    basis = np.empty((nb, Np))
    direction = np.zeros((nb, Np, 3))
    coefs = np.empty((p.shape[0], nb))
    scale = p[:, 6] if D == 7 else p[:, 9]
    for a in range(3):
        for q, pw in enumerate(powers):
            k = a * len(powers) + q
            basis[k] = i ** pw
            direction[k, :, a] = 1.0
            coefs[:, k] = p[:, a * len(powers) + q] * (scale if pw == 1 else 1.0)
    return basis, direction, coefs


def coef_grid(lo: float, hi: float, n: int, dims: int = 3) -> np.ndarray:
    """n^dims candidates on a regular grid over [lo, hi]^dims (first
    coefficient slowest)."""
    axis = np.linspace(lo, hi, n)
    return np.array(list(itertools.product(axis, repeat=dims)), dtype=np.float64)


@dataclasses.dataclass
class GridSearchResult:
    best_coef: np.ndarray
    best_entropy: float
    entropies: np.ndarray
    gpu_ms: float
    ncand: int


def grid_search(ctx, prof: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor, f0: torch.Tensor,
                grid, maxWr: float, basis: torch.Tensor, direction: torch.Tensor,
                coefs: np.ndarray, *, los=None, tile: int = 0) -> GridSearchResult:
    """Evaluate every candidate in `coefs` (ncand, nbasis) in one persistent
    launch and return the entropy minimiser.  gpu_ms is the device time of the
    launch (HIP events on the stream the work runs on)."""
    dev = prof.device
    c = torch.from_numpy(np.ascontiguousarray(coefs, dtype=np.float64)).to(dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    ent = ctx.batch_entropy(prof, ant, r0, f0, grid, maxWr, c, basis, direction, los=los,
                            tile=tile)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    ent_h = ent.cpu().numpy()
    i = int(np.argmin(ent_h))
    return GridSearchResult(best_coef=np.array(coefs[i]), best_entropy=float(ent_h[i]),
                            entropies=ent_h, gpu_ms=ms, ncand=len(coefs))


def perturbed(geom: Geometry, coef, powers=(2, 3, 4), x0: float = 0.0, y0: float = 0.0) -> Geometry:
    """Geometry whose phase centres carry the LOS polynomial error `coef`
    (R0 and frequencies unchanged: the radar's dechirp reference is kept)."""
    basis = poly_basis(geom.Np, powers)
    d = los_directions(geom.ant, x0, y0)
    return Geometry(ant=apply_los_offsets(geom.ant, coef, basis, d), r0=geom.r0.copy(),
                    f0=geom.f0.copy(), df=geom.df, K=geom.K)
