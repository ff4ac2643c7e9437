"""GPU parity of range compression and backprojection against the oracle.

Tolerances (stated per test): mixed-precision BP <= 1e-4 relative L2 against
the FP64 oracle with identical per-target peak pixels; FP32 range
compression <= 2e-6 relative L2; FP64 kernels <= 1e-10.
"""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from scene_data import c0_inputs, local_peaks, peak_index, small_inputs
from uncc_sar_focusing_amd import _lib
from uncc_sar_focusing_amd.scenes import ImageGrid, circular_aperture, point_targets

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(DEV)


def geom_dev(geom):
    return (to_dev(geom.ant, torch.float64), to_dev(geom.r0, torch.float64),
            to_dev(geom.f0, torch.float64))


# ------------------------------------------------------- range compression
@pytest.mark.parametrize("nfft", [32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384])
def test_rc_all_sizes_match_oracle(ctx, nfft):
    rng = np.random.default_rng(nfft)
    Np = 9
    for K in sorted({1, 3, min(nfft, 424), nfft // 2 + 1, nfft}):
        ph = (rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))).astype(np.complex64)
        prof = ctx.range_compress(to_dev(ph), nfft).cpu().numpy()
        ref = oracle.range_compress_f64(ph, nfft)
        e = oracle.rel_l2(prof, ref)
        assert e < 2e-6, (nfft, K, e)


def test_rc_taper_and_ramp(ctx):
    rng = np.random.default_rng(5)
    Np, K, nfft = 6, 100, 1024
    ph = (rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))).astype(np.complex64)
    taper = np.hamming(K).astype(np.float32)
    f0 = 9.6e9 + 1e6 * np.arange(Np)
    r0 = 12500.0 + 0.37 * np.arange(Np)
    prof = ctx.range_compress(to_dev(ph), nfft, taper=to_dev(taper),
                              ramp=(to_dev(f0), to_dev(r0))).cpu().numpy()
    ref = oracle.range_compress_f64(ph, nfft, taper=taper, ramp=(f0, r0))
    assert oracle.rel_l2(prof, ref) < 2e-6


@pytest.mark.parametrize("nfft", [2, 8, 1024, 16384, 32768])
def test_rc_f64_matches_oracle(ctx, nfft):
    rng = np.random.default_rng(nfft)
    Np, K = 5, max(1, nfft // 3)
    ph = rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))
    prof = ctx.range_compress(to_dev(ph), nfft).cpu().numpy()
    assert oracle.rel_l2(prof, oracle.range_compress_f64(ph, nfft)) < 1e-12


def _rc_ref(ph, nfft, taper=None, ramp=None):
    """FP64 range compression of ph: the oracle up to 16384 bins; above, the
    numpy expression the oracle is pinned to in test_oracle.py
    (roll(ifft(ph * taper, N), -((N+1)//2)), i.e. cshift((N+1)/2), times
    e^{-j 4 pi f0 r0 / c}), since the oracle's direct DFT of a length that is
    not a power of two costs O(N K)."""
    if nfft <= 16384:
        return oracle.range_compress_f64(ph, nfft, taper=taper, ramp=ramp)
    x = np.asarray(ph, np.complex128)
    if taper is not None:
        x = x * np.asarray(taper, np.float64)[None, :]
    x = np.roll(np.fft.ifft(x, nfft, axis=1), -((nfft + 1) // 2), axis=1)
    if ramp is not None:
        f0, r0 = (np.asarray(v, np.float64) for v in ramp)
        x = x * np.exp(-1j * 4 * np.pi / 299792458.0 * f0 * r0)[:, None]
    return x


@pytest.mark.parametrize("nfft", [3, 37, 424, 1000, 3000, 8191, 8192 + 57, 30011, 100003,
                                  (1 << 20) - 3])
def test_rc_non_power_of_two_matches_oracle(ctx, nfft):
    """Nfft that is not a power of two (N_fft = K is the reference's default,
    charlotte_sar_api.hpp:173): Bluestein's identity, register-resident up
    to 8192 (odd lengths included: the fftshift remap) and through the
    global-memory passes beyond (and in FP64); FP32 <= 2e-6 relative L2
    against the FP64 reference on the same complex64 input, with taper and
    ramp on one case; FP64 <= 1e-12."""
    rng = np.random.default_rng(nfft)
    Np = 5
    for K in sorted({1, min(nfft, 424), nfft}):
        ph = rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))
        p32 = ctx.range_compress(to_dev(ph.astype(np.complex64)), nfft).cpu().numpy()
        e = oracle.rel_l2(p32, _rc_ref(ph.astype(np.complex64), nfft))
        assert e < 2e-6, (nfft, K, e)
    p64 = ctx.range_compress(to_dev(ph), nfft).cpu().numpy()
    assert oracle.rel_l2(p64, _rc_ref(ph, nfft)) < 1e-12
    Kt = min(nfft, 424)
    ph32 = (rng.standard_normal((Np, Kt)) + 1j * rng.standard_normal((Np, Kt))).astype(np.complex64)
    taper = np.hanning(Kt).astype(np.float32)
    f0 = 9.6e9 + 1e6 * np.arange(Np)
    r0 = 12500.0 + 0.37 * np.arange(Np)
    pt = ctx.range_compress(to_dev(ph32), nfft, taper=to_dev(taper),
                            ramp=(to_dev(f0), to_dev(r0))).cpu().numpy()
    ref = _rc_ref(ph32, nfft, taper=taper, ramp=(f0, r0))
    assert oracle.rel_l2(pt, ref) < 2e-6


@pytest.mark.parametrize("nfft", [2, 4, 16, 32768, 65536, 1 << 18])
def test_rc_power_of_two_outside_lds_range(ctx, nfft):
    """Powers of two below 32 and above 16384 (global-memory Stockham passes
    in FP32): <= 2e-6 relative L2 against the FP64 oracle."""
    rng = np.random.default_rng(nfft)
    Np = 3
    for K in sorted({1, max(1, nfft // 3), nfft}):
        ph = (rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))).astype(np.complex64)
        prof = ctx.range_compress(to_dev(ph), nfft).cpu().numpy()
        e = oracle.rel_l2(prof, oracle.range_compress_f64(ph, nfft))
        assert e < 2e-6, (nfft, K, e)


def test_rc_rejects_bad_sizes(ctx):
    ph = to_dev(np.zeros((2, 10), np.complex64))
    with pytest.raises(_lib.SarfError) as e:
        ctx.range_compress(ph, 8)  # K > Nfft
    assert e.value.code == _lib.SARF_ERR_INVALID
    with pytest.raises(_lib.SarfError) as e:
        ctx.range_compress(ph, (1 << 20) + 1)  # beyond the Bluestein range
    assert e.value.code == _lib.SARF_ERR_UNSUPPORTED


def _rc_scene_inputs(seed):
    """Uneven scenes spanning every transform-length class the batched call
    groups by: several per Nfft, 4 and 16 sample rows per lane (K <= Nfft/4
    or not), one pulse and pulse counts that leave partial pulse groups,
    taper and ramp on some."""
    rng = np.random.default_rng(seed)
    specs = [(1, 20, 32), (7, 32, 32), (33, 5, 64), (9, 100, 256), (300, 64, 256),
             (17, 1000, 1024), (5, 200, 1024), (64, 256, 4096), (3, 4096, 4096),
             (40, 1500, 8192), (2, 300, 16384), (11, 9000, 16384)]
    out = []
    for i, (Np, K, nfft) in enumerate(specs):
        ph = (rng.standard_normal((Np, K)) + 1j * rng.standard_normal((Np, K))).astype(np.complex64)
        taper = np.hanning(K).astype(np.float32) if i % 3 == 1 else None
        ramp = (9.6e9 + 1e6 * np.arange(Np), 12500.0 + 0.37 * np.arange(Np)) if i % 4 == 2 \
            else None
        out.append((ph, nfft, taper, ramp))
    return out


def test_rc_scenes_equal_per_scene_calls(ctx):
    """sarf_range_compress_scenes: bit-identical to one sarf_range_compress
    call per scene (same kernel body), and within 2e-6 rel-L2 of the FP64
    oracle; a second call with the same list (cached upload) and a call
    with fewer scenes give the same results."""
    inp = _rc_scene_inputs(11)
    phs = [to_dev(ph) for ph, _, _, _ in inp]
    tapers = [None if t is None else to_dev(t) for _, _, t, _ in inp]
    ramps = [None if r is None else (to_dev(r[0]), to_dev(r[1])) for _, _, _, r in inp]
    nffts = [n for _, n, _, _ in inp]
    single = [ctx.range_compress(p, n, taper=t, ramp=r).cpu()
              for p, n, t, r in zip(phs, nffts, tapers, ramps)]
    outs = ctx.range_compress_scenes(phs, nffts, tapers=tapers, ramps=ramps)
    for i, (o, s1) in enumerate(zip(outs, single)):
        assert torch.equal(o.cpu(), s1), i
        ph, nfft, taper, ramp = inp[i]
        ref = oracle.range_compress_f64(ph, nfft, taper=taper, ramp=ramp)
        assert oracle.rel_l2(o.cpu().numpy(), ref) < 2e-6, i
    for o in outs:
        o.zero_()
    ctx.range_compress_scenes(phs, nffts, tapers=tapers, ramps=ramps, out=outs)
    for i, (o, s1) in enumerate(zip(outs, single)):
        assert torch.equal(o.cpu(), s1), i
    sub = [1, 4, 9]
    outs2 = ctx.range_compress_scenes([phs[i] for i in sub], [nffts[i] for i in sub],
                                      tapers=[tapers[i] for i in sub],
                                      ramps=[ramps[i] for i in sub])
    for o, i in zip(outs2, sub):
        assert torch.equal(o.cpu(), single[i]), i


def test_rc_scenes_rejects_bad_scenes(ctx):
    ph = to_dev(np.zeros((2, 10), np.complex64))
    ok = to_dev(np.zeros((2, 16), np.complex64))
    with pytest.raises(_lib.SarfError) as e:
        ctx.range_compress_scenes([ok, ph], [32, 8])  # K > Nfft in scene 1
    assert e.value.code == _lib.SARF_ERR_INVALID and "scene 1" in str(e.value)
    with pytest.raises(_lib.SarfError) as e:
        ctx.range_compress_scenes([ph], [24])  # not a power of two
    assert e.value.code == _lib.SARF_ERR_UNSUPPORTED
    with pytest.raises(_lib.SarfError) as e:
        ctx.range_compress_scenes([], [])
    assert e.value.code == _lib.SARF_ERR_INVALID


# ----------------------------------------------------------- backprojection
def test_bp_c0_full_matches_fp64_oracle(ctx):
    """Configuration C0 in full: RC + mixed BP on the GPU vs FP64 oracle.
    Tolerance: rel-L2 <= 1e-4, identical per-target and global peak pixels."""
    cfg, geom, tg, grid, ph = c0_inputs()
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), cfg.nfft)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, cfg.nfft), geom.ant, geom.r0, geom.f0,
                        grid, geom.maxWr)
    e = oracle.rel_l2(img, ref)
    print(f"C0 mixed rel-L2 vs FP64 oracle: {e:.3e}")
    assert e <= 1e-4
    assert peak_index(img) == peak_index(ref)
    assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)


def test_bp_c0_f64_matches_oracle(ctx):
    cfg, geom, tg, grid, ph = c0_inputs()
    ant, r0, f0 = geom_dev(geom)
    ph64 = to_dev(ph.astype(np.complex128))
    prof = ctx.range_compress(ph64, cfg.nfft)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, precision="f64").cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, cfg.nfft), geom.ant, geom.r0, geom.f0,
                        grid, geom.maxWr)
    assert oracle.rel_l2(img, ref) < 1e-10


@pytest.mark.parametrize("tile", [1, 2, 3, 4, 5, 6, 7, 11, 12, 13, 14, 15, 16])
def test_bp_every_tile_shape(ctx, tile):
    geom, tg, ph = small_inputs()
    grid = ImageGrid(70, 45, 0.4, -0.3, 7.0, 5.5)  # ragged vs every tile
    nfft = 1024
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), nfft)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=tile).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, nfft), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4, tile
    assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)


@pytest.mark.parametrize("K,nfft", [(96, 96), (96, 300), (424, 424)])
def test_bp_non_power_of_two_nfft_matches_oracle(ctx, K, nfft):
    """Nfft = K, the reference's default (charlotte_sar_api.hpp:173), and other
    lengths that are not powers of two: direct-DFT range compression, then
    mixed BP (every bin width and window follow Nfft), and the host-buffer
    path, against the FP64 oracle.  Tolerance: rel-L2 <= 1e-4, same peaks."""
    geom, tg, ph = small_inputs(Np=64, K=K)
    grid = ImageGrid(48, 40, 0.2, -0.1, 6.0, 5.0)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), nfft)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, nfft), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4
    assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)
    host = ctx.focus_host(ph, geom.ant, geom.r0, geom.f0, grid, nfft, geom.maxWr)
    assert oracle.rel_l2(host, ref) <= 1e-4


@pytest.mark.parametrize("Np", [1, 31, 33, 100])
def test_bp_pulse_counts(ctx, Np):
    geom, tg, ph = small_inputs(Np=Np, K=64)
    grid = ImageGrid.square(40, 0.15)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 512)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, 512), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4


@pytest.mark.parametrize("f0,tile,tol", [(35e9, 0, 1e-4), (35e9, 4, 1e-4), (35e9, 7, 1e-4),
                                         (35e9, 13, 1e-4), (94e9, 0, 3e-4)])
def test_bp_high_carrier_matches_oracle(ctx, f0, tile, tol):
    """Ka/W-band carriers: 2 f0 |dR - dR_c| / c over a tile reaches several
    hundred revolutions, past the sine unit's input range, so those pulses
    take the phase-reducing path.  FP32 phase rounding grows with f0 times
    the tile extent (1.2e-5 at 9.6 GHz on this scene), so the W-band case is
    held to 3e-4; the FP64 path is the route to 1e-4 there."""
    geom = circular_aperture(64, 96, f0=f0)
    tg = small_inputs()[1]
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    grid = ImageGrid(70, 45, 0.4, -0.3, 7.0, 5.5)
    nfft = 1024
    ant, r0, f0d = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), nfft)
    img = ctx.backproject(prof, ant, r0, f0d, grid, geom.maxWr, tile=tile).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, nfft), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= tol, (f0, tile)
    assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)
    if f0 > 50e9:
        prof64 = ctx.range_compress(to_dev(ph.astype(np.complex128)), nfft)
        img = ctx.backproject(prof64, ant, r0, f0d, grid, geom.maxWr,
                              precision="f64").cpu().numpy()
        assert oracle.rel_l2(img, ref) <= 1e-10


def test_bp_swath_edge_gate(ctx):
    """Image larger than the unambiguous range window: pixels outside
    (r_0, r_{N-1}) must stay zero as in the reference (strict gate)."""
    geom, tg, ph = small_inputs(Np=16, K=32, extent=2.0)
    # maxWr = c/(2 df) ~ 7.3 m for K = 32, 640 MHz; image 24 m down-range
    grid = ImageGrid(8, 160, 0.0, 0.0, 1.4, 24.0)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 256)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, 256), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    zero_ref = ref == 0
    assert zero_ref.sum() > 100
    # exact agreement on the gate except pixels within 1e-5 m of the edge
    mism = (img == 0) != zero_ref
    assert mism.sum() <= 2
    assert oracle.rel_l2(img, ref) <= 1e-4


def test_bp_accumulate_is_pulse_split_additive(ctx):
    geom, tg, ph = small_inputs(Np=80, K=64)
    grid = ImageGrid.square(48, 0.12)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 512)
    full = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr)
    part = ctx.backproject(prof[:30].contiguous(), ant[:30].contiguous(), r0[:30].contiguous(),
                           f0[:30].contiguous(), grid, geom.maxWr)
    ctx.backproject(prof[30:].contiguous(), ant[30:].contiguous(), r0[30:].contiguous(),
                    f0[30:].contiguous(), grid, geom.maxWr, out=part, accumulate=True)
    assert oracle.rel_l2(part.cpu().numpy(), full.cpu().numpy()) < 1e-6


def test_bp_is_deterministic(ctx):
    geom, tg, ph = small_inputs()
    grid = ImageGrid.square(64, 0.1)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 1024)
    a = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu()
    b = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu()
    assert torch.equal(a, b)


@pytest.mark.parametrize("tile", [0, 1, 4, 7, 11, 13, 15])
def test_bp_smallest_images(ctx, tile):
    """Images far smaller than a tile (2x2, 2x3, 3x2, 7x2) centred on a
    target's main lobe: the tile's pixels outside the image are neither
    computed into nor written; rel-L2 <= 1e-4 against the oracle."""
    geom, tg, ph = small_inputs(Np=20, K=64)
    ant, r0, f0 = geom_dev(geom)
    nfft = 512
    prof = ctx.range_compress(to_dev(ph), nfft)
    prof_h = oracle.range_compress_f64(ph, nfft)
    cx, cy = tg.pos[0, 0], tg.pos[0, 1]
    for nx, ny in [(2, 2), (2, 3), (3, 2), (7, 2)]:
        grid = ImageGrid(nx, ny, cx, cy, 0.03 * (nx - 1), 0.03 * (ny - 1))
        img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=tile).cpu().numpy()
        ref = oracle.bp_f64(prof_h, geom.ant, geom.r0, geom.f0, grid, geom.maxWr)
        assert img.shape == (nx, ny)
        assert oracle.rel_l2(img, ref) <= 1e-4, (tile, nx, ny)


def test_bp_rejects_bad_descriptor(ctx):
    geom, tg, ph = small_inputs()
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 1024)
    with pytest.raises(_lib.SarfError) as e:
        ctx.backproject(prof, ant, r0, f0, ImageGrid(1, 10, 0, 0, 1.0, 1.0), geom.maxWr)
    assert e.value.code == _lib.SARF_ERR_INVALID
    with pytest.raises(_lib.SarfError):
        ctx.backproject(prof, ant, r0, f0, ImageGrid(8, 8, 0, 0, 1.0, 1.0), -1.0)


def test_focus_host_matches_oracle(ctx):
    geom, tg, ph = small_inputs()
    grid = ImageGrid(50, 60, 0.0, 0.0, 6.0, 7.0)
    img = ctx.focus_host(ph, geom.ant, geom.r0, geom.f0, grid, 1024, geom.maxWr)
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, 1024), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4


@pytest.mark.parametrize("Np", [1024, 3000, 4100])
def test_focus_host_pulse_groups_match_oracle(ctx, Np):
    """Host-buffer focusing uploads pulse groups while earlier groups are
    compressed and backprojected (one group at 1024; at 3000 a first group
    of 375 pulses, then 2 of about 1312; at 4100 a first of 512, then 4 of
    about 897); the accumulated image meets the same bound."""
    geom, tg, ph = small_inputs(Np=Np, K=64, n_targets=3)
    grid = ImageGrid(40, 36, 0.2, -0.1, 4.0, 3.5)
    img = ctx.focus_host(ph, geom.ant, geom.r0, geom.f0, grid, 256, geom.maxWr)
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, 256), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4
    assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)


def test_focus_host_page_locked_buffers_match_pageable(ctx):
    """Page-locked phase history and image (torch tensors made with
    pin_memory=True) are copied to and from the GPU without the pinned
    staging; every combination gives exactly the image of the pageable call
    (same kernels, same pulse groups: 3000 pulses = 375 + 2 x ~1312)."""
    geom, tg, ph = small_inputs(Np=3000, K=64, n_targets=3)
    grid = ImageGrid(40, 36, 0.2, -0.1, 4.0, 3.5)
    args = (geom.ant, geom.r0, geom.f0, grid, 256, geom.maxWr)
    ref = ctx.focus_host(ph, *args)
    ph_pin = torch.empty(ph.shape, dtype=torch.complex64, pin_memory=True)
    ph_pin.copy_(torch.from_numpy(ph))
    for pin_in in (False, True):
        for pin_out in (False, True):
            out = (torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64, pin_memory=True).numpy()
                   if pin_out else np.empty((grid.Nx, grid.Ny), dtype=np.complex64))
            img = ctx.focus_host(ph_pin.numpy() if pin_in else ph, *args, out=out)
            assert img is out
            assert np.array_equal(img, ref), (pin_in, pin_out)


def test_synth_matches_oracle(ctx):
    geom = circular_aperture(33, 300)
    from uncc_sar_focusing_amd.scenes import point_targets
    tg = point_targets(300, 30.0, seed=1)
    ant, r0, f0 = geom_dev(geom)
    ph = ctx.synth_phase_history(ant, r0, f0, geom.df, geom.K, to_dev(tg.pos),
                                 to_dev(tg.amp.astype(np.complex64))).cpu().numpy()
    ref = oracle.synth_ph_f64(geom, tg)
    assert oracle.rel_l2(ph, ref) < 2e-5


@pytest.mark.parametrize("split", [1, 3, 8, 64])
def test_bp_pulse_split_partials_match_oracle(ctx, split):
    """Pulse groups forming partial images summed in order (the path used
    when an image has too few tiles to fill the GPU)."""
    geom, tg, ph = small_inputs(Np=100, K=64)
    grid = ImageGrid.square(40, 0.15)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), 512)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, pulse_split=split).cpu().numpy()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, 512), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    assert oracle.rel_l2(img, ref) <= 1e-4
    # accumulate on top of an existing image
    base = to_dev(np.full((grid.Nx, grid.Ny), 1 + 2j, np.complex64))
    ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, pulse_split=split, out=base,
                    accumulate=True)
    assert oracle.rel_l2(base.cpu().numpy() - (1 + 2j), ref) <= 1e-4


@pytest.mark.parametrize("split", [0, 2, 5])
def test_bp_head_tail_split_matches_oracle(ctx, split):
    """An image whose tiles fill two rounds of resident workgroups (2048^2
    with tile 13: 2048 tiles): one head group of the first 7/8 of the pulses
    writes the image directly and split - 1 tail groups add partial images
    (0 = the automatic count, 3).  Within 1e-4 rel-L2 of the FP64 oracle
    with the same brightest pixel, also when accumulating onto an existing
    image, and within 2e-6 of the unsplit image."""
    Np, K, n, sp, nfft = 128, 256, 2048, 0.05, 2048
    geom = circular_aperture(Np, K)
    tg = point_targets(5, 0.6 * sp * n, seed=9)
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    grid = ImageGrid.square(n, sp)
    ant, r0, f0 = geom_dev(geom)
    prof = ctx.range_compress(to_dev(ph), nfft)
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=13, pulse_split=split)
    one = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=13, pulse_split=1)
    got = img.cpu().numpy()
    assert oracle.rel_l2(got, one.cpu().numpy()) <= 2e-6
    ref = oracle.bp_f64(oracle.range_compress_f64(ph, nfft), geom.ant, geom.r0, geom.f0, grid,
                        geom.maxWr)
    e = oracle.rel_l2(got, ref)
    print(f"head+tail split={split}: rel-L2 vs oracle {e:.3e}", flush=True)
    assert e <= 1e-4
    assert peak_index(got) == peak_index(ref)
    base = to_dev(np.full((grid.Nx, grid.Ny), 1 + 2j, np.complex64))
    ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=13, pulse_split=split, out=base,
                    accumulate=True)
    assert oracle.rel_l2(base.cpu().numpy() - (1 + 2j), ref) <= 1e-4


def test_bp_scenes_uneven_batch_matches_single_calls(ctx):
    """One persistent launch over scenes of different pulses, profile
    lengths and image sizes equals backprojecting each scene alone, and the
    FP64 oracle within 1e-4.  A mixed batch (a tile per scene) is refused
    here: at these coarse spacings none of its tiles 15, 16, 13 meets the
    accuracy bound (the full-size C4 test runs it)."""
    specs = [(40, 64, 256, ImageGrid(37, 52, 0.1, -0.2, 4.0, 5.0)),
             (97, 96, 1024, ImageGrid.square(64, 0.1)),
             (16, 48, 512, ImageGrid(20, 90, 0.0, 0.3, 2.0, 9.0))]
    scenes, refs = [], []
    for i, (Np, K, nfft, grid) in enumerate(specs):
        geom, tg, ph = small_inputs(Np=Np, K=K, seed=10 + i)
        ant, r0, f0 = geom_dev(geom)
        prof = ctx.range_compress(to_dev(ph), nfft)
        scenes.append((prof, ant, r0, f0, grid, geom.maxWr))
        refs.append((ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu().numpy(),
                     oracle.bp_f64(oracle.range_compress_f64(ph, nfft), geom.ant, geom.r0,
                                   geom.f0, grid, geom.maxWr)))
    for tile in (0, 1, 5, 11, 13, 15, 16):
        imgs = ctx.backproject_scenes(scenes, tile=tile)
        for img, (single, ref) in zip(imgs, refs):
            img = img.cpu().numpy()
            assert oracle.rel_l2(img, ref) <= 1e-4
            if tile == 0:
                assert oracle.rel_l2(img, single) < 1e-5
    with pytest.raises(_lib.SarfError) as e:
        ctx.backproject_scenes(scenes, tile=_lib.SARF_SCENES_TILE_MIXED)
    assert e.value.code == _lib.SARF_ERR_INVALID
    # the batch tuner returns an accurate tile whose batch meets the same bound
    # (not the mixed mode, which these scenes cannot take)
    best, ms = ctx.autotune_scenes(scenes, reps=1)
    assert best in (1, 2, 3, 4, 5, 11, 12, 13, 14, 15, 16) and ms > 0
    imgs = ctx.backproject_scenes(scenes, tile=best)
    for img, (single, ref) in zip(imgs, refs):
        assert oracle.rel_l2(img.cpu().numpy(), ref) <= 1e-4


@pytest.mark.parametrize("K,Np,S", [(300, 17, 2000), (64, 9, 500), (1025, 5, 3000)])
def test_synth_nufft_matches_direct_oracle(ctx, K, Np, S):
    """Non-uniform-FFT clutter synthesis vs the exact FP64 sum; includes
    scatterers beyond the unambiguous range (periodic wrap).  <= 2e-5."""
    from uncc_sar_focusing_amd.scenes import Targets, uniform
    geom = circular_aperture(Np, K)
    u = uniform(77, 3, 3 * S).reshape(S, 3)
    pos = np.zeros((S, 3))
    pos[:, 0] = (u[:, 0] - 0.5) * 30.0
    pos[:, 1] = (u[:, 1] - 0.5) * 1.5 * geom.maxWr
    amp = np.exp(2j * np.pi * u[:, 2]) * (0.1 + u[:, 0])
    tg = Targets(pos, amp)
    ant, r0, f0 = geom_dev(geom)
    ph = ctx.synth_phase_history(ant, r0, f0, geom.df, K, to_dev(pos),
                                 to_dev(amp.astype(np.complex64)), method="nufft").cpu().numpy()
    ref = oracle.synth_ph_f64(geom, tg)
    assert oracle.rel_l2(ph, ref) < 2e-5, oracle.rel_l2(ph, ref)


def test_autotune_chooses_an_accurate_shape_and_is_used_by_default():
    """sarf_autotune on C0: the chosen (tile, split) image meets the same
    1e-4 bound as the default; a later default-shaped call on the same
    geometry reuses it (bit-identical to the explicit pair); other
    geometries are unaffected."""
    from uncc_sar_focusing_amd.api import Context

    tctx = Context(0)  # own context: the tuned shape is per context
    try:
        cfg, geom, tg, grid, ph = c0_inputs()
        ant, r0, f0 = geom_dev(geom)
        prof = tctx.range_compress(to_dev(ph), cfg.nfft)
        res, img = tctx.autotune(prof, ant, r0, f0, grid, geom.maxWr, reps=2)
        assert res["tile"] in (1, 2, 3, 4, 5, 6, 7, 11, 12, 13, 14, 15, 16) and res["pulse_split"] >= 1
        assert res["candidates"] >= 2
        assert 0 < res["ms"] <= res["default_ms"] * 1.0001
        ref = oracle.bp_f64(oracle.range_compress_f64(ph, cfg.nfft), geom.ant, geom.r0, geom.f0,
                            grid, geom.maxWr)
        img = img.cpu().numpy()
        assert oracle.rel_l2(img, ref) <= 1e-4
        assert local_peaks(img, grid, tg.pos) == local_peaks(ref, grid, tg.pos)
        dflt = tctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu()
        expl = tctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=res["tile"],
                                pulse_split=res["pulse_split"]).cpu()
        assert torch.equal(dflt, expl)
        assert torch.equal(dflt, torch.from_numpy(img))
        with pytest.raises(_lib.SarfError) as e:
            tctx.autotune(prof, ant, r0, f0, grid, geom.maxWr, reps=0)
        assert e.value.code == _lib.SARF_ERR_INVALID
    finally:
        tctx.close()


def test_tune_table_round_trips_through_json_into_another_context():
    """The remembered shapes export as a JSON-ready table; loaded into a new
    context they steer its default-shaped calls on the same geometry (here
    a pair other than the default, so the effect is visible: the default
    call equals that explicit pair bit for bit and not the untuned default),
    and a table with an unknown tile is refused whole."""
    import json

    from uncc_sar_focusing_amd.api import Context

    cfg, geom, tg, grid, ph = c0_inputs()
    a_ctx, b_ctx = Context(0), Context(0)
    try:
        ant, r0, f0 = geom_dev(geom)
        prof = a_ctx.range_compress(to_dev(ph), cfg.nfft)
        untuned = b_ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu()
        res, _ = a_ctx.autotune(prof, ant, r0, f0, grid, geom.maxWr, reps=1)
        table = json.loads(json.dumps(a_ctx.tune_table()))
        assert len(table) == 1
        e = table[0]
        assert (e["Nx"], e["Ny"], e["Np"], e["Nfft"]) == (grid.Nx, grid.Ny, geom.Np, cfg.nfft)
        assert (e["tile"], e["pulse_split"]) == (res["tile"], res["pulse_split"])
        # a pair that differs from the untuned default (tile 11 x 2 splits, or 1 x 2)
        e["tile"], e["pulse_split"] = (1, 2) if res["tile"] == 11 else (11, 2)
        bad = dict(e, tile=99)
        with pytest.raises(_lib.SarfError) as err:
            b_ctx.load_tune_table([e, bad])
        assert err.value.code == _lib.SARF_ERR_INVALID
        assert b_ctx.tune_table() == []
        b_ctx.load_tune_table([e])
        assert b_ctx.tune_table() == [e]
        dflt = b_ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr).cpu()
        expl = b_ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, tile=e["tile"],
                                 pulse_split=e["pulse_split"]).cpu()
        assert torch.equal(dflt, expl)
        assert not torch.equal(dflt, untuned)
        ref = oracle.bp_f64(oracle.range_compress_f64(ph, cfg.nfft), geom.ant, geom.r0, geom.f0,
                            grid, geom.maxWr)
        assert oracle.rel_l2(dflt.numpy(), ref) <= 1e-4
    finally:
        a_ctx.close()
        b_ctx.close()


# ------------------------------------------------------------ matched filter
@pytest.mark.parametrize("Np,K,Nx,Ny,split", [(24, 96, 40, 36, 0), (8, 1501, 20, 70, 0),
                                              (33, 64, 17, 130, 5), (3, 1, 5, 3, 0)])
def test_matched_filter_matches_fp64_oracle(ctx, Np, K, Nx, Ny, split):
    """Exact matched filter (run_mf / mfBasic.m) on the GPU vs the FP64
    oracle, including K over one LDS stage with an odd tail, ragged tiles,
    explicit pulse splits and K = 1.  Tolerance: rel-L2 <= 1e-4."""
    geom, tg, ph = small_inputs(Np=Np, K=K, extent=2.0)
    grid = ImageGrid(Nx, Ny, 0.2, -0.1, 3.0, 2.5)
    ant, r0, f0 = geom_dev(geom)
    img = ctx.matched_filter(to_dev(ph), ant, r0, f0, geom.df, grid, pulse_split=split)
    ref = oracle.mf_f64(ph, geom.ant, geom.r0, geom.f0, geom.df, grid)
    e = oracle.rel_l2(img.cpu().numpy(), ref)
    print(f"MF Np={Np} K={K}: rel-L2 {e:.2e}")
    assert e <= 1e-4


def test_matched_filter_accumulate_and_limits(ctx):
    geom, tg, ph = small_inputs(Np=20, K=64, extent=2.0)
    grid = ImageGrid.square(24, 0.1)
    ant, r0, f0 = geom_dev(geom)
    phd = to_dev(ph)
    full = ctx.matched_filter(phd, ant, r0, f0, geom.df, grid)
    part = ctx.matched_filter(phd[:7].contiguous(), ant[:7].contiguous(), r0[:7].contiguous(),
                              f0[:7].contiguous(), geom.df, grid)
    ctx.matched_filter(phd[7:].contiguous(), ant[7:].contiguous(), r0[7:].contiguous(),
                       f0[7:].contiguous(), geom.df, grid, out=part, accumulate=True)
    assert oracle.rel_l2(part.cpu().numpy(), full.cpu().numpy()) < 1e-6
    big = torch.zeros((1, 8193), dtype=torch.complex64, device=DEV)
    with pytest.raises(_lib.SarfError) as e:
        ctx.matched_filter(big, ant[:1].contiguous(), r0[:1].contiguous(), f0[:1].contiguous(),
                           geom.df, grid)
    assert e.value.code == _lib.SARF_ERR_UNSUPPORTED


# --------------------------------------------------------------- dB output
def _levels(dyn):
    lv = np.zeros(256, np.float32)
    _lib.check(_lib.lib().sarf_db8_levels(dyn, lv.ctypes.data_as(ctypes.c_void_p)))
    return lv


@pytest.mark.parametrize("Nx,Ny,dyn", [(256, 256, 70.0), (70, 45, 70.0), (1, 130, 70.0),
                                       (333, 77, 45.5), (2048, 2048, 70.0)])
def test_image_db8_matches_oracle(ctx, Nx, Ny, dyn):
    """GPU 8-bit dB display image vs the oracle's restatement of the
    reference's float quantisation (host log10f): every byte equal, row
    order and the maximum included.  The 2048 x 2048 case is a C1-sized
    image of log-uniform magnitudes (1e-5 .. 1) that populates every level;
    the others are backprojected scenes or Gaussian noise."""
    rng = np.random.default_rng(9 + Nx)
    if Nx == 2048 or Nx == 1:
        mag = 10.0 ** rng.uniform(-5.0, 0.0, (Nx, Ny)) if Nx == 2048 else \
            np.abs(rng.standard_normal((Nx, Ny)))
        ph = np.exp(2j * np.pi * rng.uniform(size=(Nx, Ny)))
        img = to_dev((mag * ph).astype(np.complex64))
    else:
        geom, tg, ph = small_inputs()
        grid = ImageGrid(Nx, Ny, 0.1, -0.2, 0.05 * (Nx - 1), 0.05 * (Ny - 1))
        ant, r0, f0 = geom_dev(geom)
        prof = ctx.range_compress(to_dev(ph), 1024)
        img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr)
    got = ctx.image_db8(img, dyn).cpu().numpy()
    ref = oracle.db8(img.cpu().numpy(), dyn)
    assert got.shape == ref.shape == (Ny, Nx)
    d = np.abs(got.astype(int) - ref.astype(int))
    print(f"db8 {Nx}x{Ny} dyn {dyn}: {np.count_nonzero(d)} bytes differ, levels used "
          f"{len(np.unique(ref))}")
    assert d.max() == 0
    # the brightest pixel (r = 1) gets the expression's top byte: 255 at
    # 70 dB, 254 at 45.5 dB, where (float)(255/dyn) * dyn rounds below 255
    assert got.max() == ref.max() >= 254


@pytest.mark.parametrize("dyn", [70.0, 40.0])
def test_image_db8_at_every_level_threshold(ctx, dyn):
    """Pixels placed exactly on each level threshold and one float below it
    (max |I| = 1, real values, so r = |I| exactly): the GPU bytes equal the
    oracle's on every one -- the values where a device log10 can land on
    the other side of an integer."""
    lv = _levels(dyn)
    finite = lv[1:][np.isfinite(lv[1:])]
    vals = np.concatenate([[1.0, 0.0], finite, np.nextafter(finite, np.float32(0)),
                           np.nextafter(finite, np.float32(2))]).astype(np.float32)
    n = vals.size
    Ny = 64
    Nx = (n + Ny - 1) // Ny
    img = np.zeros(Nx * Ny, dtype=np.complex64)
    img[:n] = vals
    img = img.reshape(Nx, Ny)
    got = ctx.image_db8(to_dev(img), dyn).cpu().numpy()
    ref = oracle.db8(img, dyn)
    assert np.array_equal(got, ref)
    assert len(np.unique(ref)) >= 250


def test_image_db8_of_a_zero_image_is_zero(ctx):
    img = torch.zeros((40, 30), dtype=torch.complex64, device=DEV)
    got = ctx.image_db8(img, 70.0).cpu().numpy()
    assert got.shape == (30, 40) and got.max() == 0
