"""Seeded synthetic SAR scenes (geometry, targets, noise).

Data model (matlab/formImagePointTargetBP.m:50-81):
  * frequencies f_k = f0 + k*df, k = 0..K-1 (f0 = start frequency, the
    reference's startF / minF, uncc_sar_focusing.hpp:275; bpBasic.m:90);
  * a circular aperture: antenna a_n = (G cos th_n, G sin th_n, h),
    th_n = linspace(th_c - span/2, th_c + span/2, Np); th_c = 90 deg makes
    y down-range and x cross-range;
  * R0_n = |a_n| (the reference's default slant range,
    uncc_sar_focusing.hpp:312-323), scene at z = 0;
  * ph[n,k] = sum_t A_t exp(-j 4 pi f_k (|a_n - p_t| - R0_n)/c) + noise.

Random numbers come from a counter-based splitmix64 stream so that any
implementation (numpy here, C or HIP elsewhere) can reproduce a scene from
its seed alone.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

CLIGHT = 299792458.0
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n doubles in [0, 1) from stream `stream` of generator `seed`."""
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLD + np.uint64(stream))
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = _mix64(key + i * _GOLD)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normal doubles (Box-Muller on two uniform streams)."""
    u1 = uniform(seed, 2 * stream + 1000, n)
    u2 = uniform(seed, 2 * stream + 1001, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


@dataclasses.dataclass
class Geometry:
    """Per-pulse collection geometry (the SAR_Aperture fields BP uses)."""

    ant: np.ndarray  # (Np, 3) float64 antenna phase centres (m)
    r0: np.ndarray   # (Np,) float64 slant range to scene centre (m)
    f0: np.ndarray   # (Np,) float64 start frequency (Hz)
    df: float        # frequency step (Hz)
    K: int           # samples per pulse

    @property
    def Np(self) -> int:
        return self.ant.shape[0]

    def los(self, x0: float = 0.0, y0: float = 0.0) -> tuple[float, float]:
        """Horizontal components of the unit line of sight from (x0, y0, 0)
        to the aperture's mean phase centre."""
        u = self.ant.mean(axis=0) - np.array([x0, y0, 0.0])
        u = u / np.linalg.norm(u)
        return float(u[0]), float(u[1])

    @property
    def maxWr(self) -> float:
        """Unambiguous range extent c/(2 df) (uncc_sar_focusing.hpp:430)."""
        return CLIGHT / (2.0 * self.df)


def circular_aperture(Np: int, K: int, f0: float = 9.6e9, bw: float = 640e6,
                      ground_range: float = 10e3, altitude: float = 7.5e3,
                      span_deg: float = 4.0, center_deg: float = 90.0) -> Geometry:
    th = np.deg2rad(np.linspace(center_deg - span_deg / 2, center_deg + span_deg / 2, Np))
    ant = np.stack([ground_range * np.cos(th), ground_range * np.sin(th),
                    np.full(Np, float(altitude))], axis=1)
    r0 = np.sqrt((ant ** 2).sum(axis=1))
    df = bw / (K - 1) if K > 1 else bw
    return Geometry(ant=ant, r0=r0, f0=np.full(Np, float(f0)), df=float(df), K=K)


@dataclasses.dataclass
class ImageGrid:
    """Pixel grid of SAR_ImageFormationParameters (N_x_pix, N_y_pix, x0_m,
    y0_m, Wx_m, Wy_m); pixel spacing is W/(N-1)."""

    Nx: int
    Ny: int
    x0: float = 0.0
    y0: float = 0.0
    Wx: float = 25.5
    Wy: float = 25.5

    @classmethod
    def square(cls, n: int, spacing: float, x0: float = 0.0, y0: float = 0.0) -> "ImageGrid":
        w = spacing * (n - 1)
        return cls(n, n, x0, y0, w, w)

    def x(self) -> np.ndarray:
        return self.x0 - self.Wx / 2 + np.arange(self.Nx) * (self.Wx / (self.Nx - 1))

    def y(self) -> np.ndarray:
        return self.y0 - self.Wy / 2 + np.arange(self.Ny) * (self.Wy / (self.Ny - 1))


@dataclasses.dataclass
class Targets:
    pos: np.ndarray  # (S, 3) float64
    amp: np.ndarray  # (S,) complex128


def point_targets(n: int, extent: float, seed: int, stream: int = 1) -> Targets:
    """n targets uniform in [-extent/2, extent/2]^2 at z = 0, amplitude
    U(0.5, 1) and phase U(0, 2 pi)."""
    u = uniform(seed, stream, 4 * n).reshape(n, 4)
    pos = np.zeros((n, 3))
    pos[:, 0] = (u[:, 0] - 0.5) * extent
    pos[:, 1] = (u[:, 1] - 0.5) * extent
    amp = (0.5 + 0.5 * u[:, 2]) * np.exp(2j * math.pi * u[:, 3])
    return Targets(pos=pos, amp=amp)


def grid_clutter(grid: ImageGrid, stride: int, power_db: float, seed: int,
                 stream: int = 7) -> Targets:
    """i.i.d. CN(0, 10^(power_db/10)) scatterers on every `stride`-th grid
    cell in x and y."""
    xs = grid.x()[::stride]
    ys = grid.y()[::stride]
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    n = X.size
    sig = math.sqrt(10.0 ** (power_db / 10.0) / 2.0)
    re = normal(seed, stream, n) * sig
    im = normal(seed, stream + 1, n) * sig
    pos = np.stack([X.ravel(), Y.ravel(), np.zeros(n)], axis=1)
    return Targets(pos=pos, amp=re + 1j * im)


def add_noise(ph: np.ndarray, snr_db: float, seed: int, stream: int = 3) -> np.ndarray:
    """Add CN(0, sigma^2) noise with sigma^2 = mean|ph|^2 / 10^(snr/10)."""
    p = float(np.mean(np.abs(ph) ** 2))
    sig = math.sqrt(p / 10.0 ** (snr_db / 10.0) / 2.0)
    n = ph.size
    noise = (normal(seed, stream, n) + 1j * normal(seed, stream + 1, n)) * sig
    return ph + noise.reshape(ph.shape).astype(ph.dtype)


@dataclasses.dataclass
class SceneConfig:
    name: str
    Np: int
    K: int
    nfft: int
    n_pix: int
    spacing: float
    n_targets: int
    target_extent: float
    snr_db: float | None
    clutter_stride: int = 0
    clutter_db: float = -20.0
    seed: int = 1234

    def geometry(self) -> Geometry:
        return circular_aperture(self.Np, self.K)

    def grid(self) -> ImageGrid:
        return ImageGrid.square(self.n_pix, self.spacing)

    def targets(self) -> Targets:
        return point_targets(self.n_targets, self.target_extent, self.seed)

    @property
    def pixel_pulses(self) -> int:
        return self.n_pix * self.n_pix * self.Np


# The benchmark configurations (sizes from the task's workload list).
CONFIGS = {
    # plain method, CPU-runnable
    "C0": SceneConfig("C0", Np=512, K=424, nfft=4096, n_pix=256, spacing=0.1,
                      n_targets=5, target_extent=20.0, snr_db=20.0),
    # mixed-precision BP at scale
    "C1": SceneConfig("C1", Np=4096, K=2048, nfft=16384, n_pix=2048, spacing=0.05,
                      n_targets=100, target_extent=100.0, snr_db=None,
                      clutter_stride=4, clutter_db=-20.0),
    # batched grid-search autofocus scene
    "C2": SceneConfig("C2", Np=1024, K=1024, nfft=4096, n_pix=512, spacing=0.05,
                      n_targets=20, target_extent=25.0, snr_db=None,
                      clutter_stride=4, clutter_db=-20.0),
}
