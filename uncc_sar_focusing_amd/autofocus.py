"""Autofocus drivers on top of libsarf: batched grid search over trajectory
candidates (all candidate images of a level in one persistent launch, scored
by image entropy), multi-resolution refinement with a covariance estimate,
the trajectory fits that centre the reference's grid, and the
line-of-sight (LOS) polynomial trajectory model of the batched workload.

Reference counterparts: the cuGridSearch-driven search of
src/autofocus/autofocus_gpu.cu:707-1260 -- level loop with the box shrunk
around each level's argmin (:886-948, :1036-1102; sarf_grid_search), the
bestFit / quadFit track fits (:315-416; sarf_fit_linear / sarf_fit_quadratic),
the candidate trajectory model gridSearchErrorFunctions_dft.cuh:279-291 --
with the objective the entropy E = -sum p ln p.
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


def detrended_basis(basis: np.ndarray) -> np.ndarray:
    """Each row of `basis` (nb, Np) minus its least-squares line a + b t
    over t = linspace(-1, 1, Np).  A LOS offset that is constant or linear in
    slow time only shifts the image (range and cross-range); with these
    bases a candidate changes the focus and never the image registration,
    so candidates are compared on the same scene content."""
    basis = np.asarray(basis, dtype=np.float64)
    t = np.linspace(-1.0, 1.0, basis.shape[1])
    A = np.stack([np.ones_like(t), t], axis=1)
    coef, *_ = np.linalg.lstsq(A, basis.T, rcond=None)
    return basis - (A @ coef).T


def fit_linear(values, skip: int = 1) -> np.ndarray:
    """The reference's bestFit (autofocus_gpu.cu:315-371) of a recorded
    antenna track: float32 [intercept, slope] against the pulse index."""
    from . import _lib

    v = np.ascontiguousarray(values, dtype=np.float32)
    out = np.zeros(2, dtype=np.float32)
    _lib.check(_lib.lib().sarf_fit_linear(v.ctypes.data, v.size, skip, out.ctypes.data))
    return out


def fit_quadratic(values, skip: int = 1) -> np.ndarray:
    """The reference's quadFit (autofocus_gpu.cu:373-416): float32
    [constant, linear, quadratic] against the pulse index."""
    from . import _lib

    v = np.ascontiguousarray(values, dtype=np.float32)
    out = np.zeros(3, dtype=np.float32)
    _lib.check(_lib.lib().sarf_fit_quadratic(v.ctypes.data, v.size, skip, out.ctypes.data))
    return out


def grid_covariance(start, end, num, values, best: int) -> np.ndarray:
    """The covariance estimate sarf_grid_search reports, for objective
    `values` (one per box point, first axis slowest) with minimiser index
    `best` (host only)."""
    from . import _lib

    start = np.asarray(start, dtype=np.float64)
    D = start.size
    g = _lib.GridDesc()
    g.dims, g.model, g.levels, g.shrink = D, _lib.SARF_GRID_MODEL_COEF, 1, 0
    for d in range(D):
        g.start[d], g.end[d], g.num[d] = float(start[d]), float(end[d]), int(num[d])
    v = np.ascontiguousarray(values, dtype=np.float64)
    cov = np.zeros((D, D))
    import ctypes as C

    _lib.check(_lib.lib().sarf_grid_covariance(C.byref(g), v.ctypes.data, int(best),
                                               cov.ctypes.data))
    return cov


def reference_grid_box(ant: np.ndarray, D: int, skip: int = 1):
    """The first-level box of the reference's driver around the fitted
    antenna track (autofocus_gpu.cu:842-881 for D = 7: slopes +-2.6 with 23
    samples, velocity scale 0.9..1.1 with 5; :1000-1029 for D = 10: slopes
    +-1.3 and quadratic terms +-0.65 with 5 samples each, scale 0.9..1.1
    with 5; intercepts fixed at the fit).  Returns (start, end, num)."""
    ant = np.asarray(ant, dtype=np.float64)
    Np = ant.shape[0]
    if D == 7:
        diff, n = 2.6, 23
        fits = [fit_linear(ant[:, a].astype(np.float32), skip) for a in range(3)]
        start, end, num = [], [], []
        for c in fits:
            start += [float(c[0]), float(c[1]) - diff]
            end += [float(c[0]), float(c[1]) + diff]
            num += [1, n]
    elif D == 10:
        diff, n = 1.3, 5
        fits = [fit_quadratic(ant[:, a].astype(np.float32), skip) for a in range(3)]
        start, end, num = [], [], []
        for c in fits:
            start += [float(c[0]), float(c[1]) - diff, float(c[2]) - diff / 2]
            end += [float(c[0]), float(c[1]) + diff, float(c[2]) + diff / 2]
            num += [1, n, n]
    else:
        raise ValueError("D must be 7 or 10")
    del Np
    start.append(0.9)
    end.append(1.1)
    num.append(5)
    return np.array(start), np.array(end), np.array(num)


@dataclasses.dataclass
class MultiResResult:
    best: np.ndarray          # minimiser of the last level
    best_entropy: float
    covariance: np.ndarray    # estimate at the last level (sarf_grid_covariance)
    levels: list              # per level: start, end, best, entropy, gpu_ms, ncand

    @property
    def gpu_ms(self) -> float:
        return float(sum(L["gpu_ms"] for L in self.levels))


def grid_search_multires(ctx, prof, r0, f0, grid, maxWr, start, end, num, levels: int, *,
                         model: str = "coef", ant=None, basis=None, direction=None,
                         shrink: str = "reference", los=None, tile: int = 0) -> MultiResResult:
    """Multi-resolution grid search (sarf_grid_search): `levels` boxes, each
    shrunk to half its span around the previous level's minimiser."""
    best, e, cov, lv = ctx.grid_search(prof, r0, f0, grid, maxWr, start, end, num, levels,
                                       model=model, ant=ant, basis=basis, direction=direction,
                                       shrink=shrink, tile=tile, los=los)
    return MultiResResult(best=best, best_entropy=e, covariance=cov, levels=lv)
