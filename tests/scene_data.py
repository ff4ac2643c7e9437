"""Shared seeded test scenes: inputs built on the CPU with the oracle's FP64
phase-history synthesis, so the GPU path and the oracle consume the same
complex64 samples."""
from __future__ import annotations

import functools

import numpy as np

import oracle
from uncc_sar_focusing_amd.scenes import (
    CONFIGS,
    ImageGrid,
    Targets,
    add_noise,
    circular_aperture,
    point_targets,
)


@functools.lru_cache(maxsize=None)
def c0_inputs():
    """Configuration C0 in full: 512 pulses x 424 samples, Nfft 4096,
    256 x 256 at 0.1 m, 5 targets, 20 dB SNR."""
    cfg = CONFIGS["C0"]
    geom = cfg.geometry()
    tg = cfg.targets()
    ph = oracle.synth_ph_f64(geom, tg)
    ph = add_noise(ph, cfg.snr_db, cfg.seed).astype(np.complex64)
    return cfg, geom, tg, cfg.grid(), ph


@functools.lru_cache(maxsize=None)
def small_inputs(Np=64, K=96, n_targets=4, extent=6.0, seed=3, clutter=False):
    geom = circular_aperture(Np, K)
    tg = point_targets(n_targets, extent, seed)
    if clutter:
        g = ImageGrid.square(16, extent / 15)
        from uncc_sar_focusing_amd.scenes import grid_clutter
        cl = grid_clutter(g, 1, -20.0, seed)
        tg = Targets(np.concatenate([tg.pos, cl.pos]), np.concatenate([tg.amp, cl.amp]))
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    return geom, tg, ph


def peak_index(img) -> tuple[int, int]:
    return np.unravel_index(int(np.argmax(np.abs(img))), img.shape)


def local_peaks(img, grid, pos, radius=0.5):
    """Pixel index of the brightest pixel within `radius` m of each target."""
    x, y = grid.x(), grid.y()
    out = []
    for p in pos:
        ix = np.nonzero(np.abs(x - p[0]) <= radius)[0]
        iy = np.nonzero(np.abs(y - p[1]) <= radius)[0]
        if ix.size == 0 or iy.size == 0:
            continue
        sub = np.abs(img[np.ix_(ix, iy)])
        a, b = np.unravel_index(int(np.argmax(sub)), sub.shape)
        out.append((int(ix[a]), int(iy[b])))
    return out
