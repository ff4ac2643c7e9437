"""GPU PGA against the numpy restatement of matlab/PGA.m (oracle.pga_f64) on
the same complex64 input.  Tolerances: accumulated phase within 2e-3 rad,
per-iteration rms within 1e-3, corrected image within 1e-3 relative L2
(FP32 Bluestein FFTs vs FP64 FFTs through several data-dependent
iterations)."""
import numpy as np
import pytest
import torch

import oracle
from uncc_sar_focusing_amd.scenes import ImageGrid, circular_aperture, point_targets

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _degraded_image(nx, ny, seed=4, amp=3.0):
    geom = circular_aperture(128, 128)
    tg = point_targets(6, 10.0, seed=seed)
    prof = oracle.range_compress_f64(oracle.synth_ph_f64(geom, tg), 2048)
    grid = ImageGrid(nx, ny, 0.0, 0.0, 0.2 * (nx - 1), 0.2 * (ny - 1))
    img = oracle.bp_f64(prof, geom.ant, geom.r0, geom.f0, grid, geom.maxWr)
    x = np.linspace(-1, 1, nx)
    err = amp * x ** 2 - 0.6 * amp * x ** 3 + 0.5 * amp * x ** 4
    err = err - np.polyval(np.polyfit(np.arange(nx), err, 1), np.arange(nx))
    bad = np.fft.fft(np.fft.ifft(img, axis=0) * np.exp(-1j * err)[:, None], axis=0)
    return img, bad.astype(np.complex64)


# transform lengths M = 2^L >= 2 nx - 1: L = 7..9 and 5 take the in-LDS
# row kernels, L = 10..13 the register-resident ones
@pytest.mark.parametrize("nx,ny", [(64, 64), (75, 50), (128, 97), (333, 40), (1024, 31),
                                   (4096, 6), (2, 9), (600, 96), (700, 300), (1500, 16)])
def test_pga_matches_oracle(ctx, nx, ny):
    img, bad = _degraded_image(nx, ny)
    iters = 3
    t = torch.from_numpy(bad.copy()).to(DEV)
    af_ph, rms = ctx.pga(t, iters)
    out = t.cpu().numpy()
    ref_img, ref_ph, ref_rms = oracle.pga_f64(bad, iters)
    assert np.max(np.abs(af_ph.cpu().numpy() - ref_ph)) < 2e-3
    assert np.max(np.abs(rms.cpu().numpy() - np.array(ref_rms))) < 1e-3
    assert oracle.rel_l2(out, ref_img) < 1e-3
    assert abs(oracle.entropy_f64(out) - oracle.entropy_f64(ref_img)) < 1e-3
    if 64 <= nx <= 128:  # square-ish scenes: PGA focuses them (degenerate ones need not)
        assert oracle.entropy_f64(out) < oracle.entropy_f64(bad) - 0.3


def test_pga_zero_iterations_is_identity(ctx):
    img, bad = _degraded_image(64, 64)
    t = torch.from_numpy(bad.copy()).to(DEV)
    af_ph, rms = ctx.pga(t, 0)
    assert torch.equal(t.cpu(), torch.from_numpy(bad))
    assert float(af_ph.abs().max()) == 0.0


def test_pga_scenes_equal_per_image_calls(ctx):
    """sarf_pga_scenes on images of every transform length (M = 32 ... 8192,
    both row-kernel families, uneven heights) gives exactly what sarf_pga
    gives on each image alone; a second call with the same list reuses the
    uploaded state and gives the same result again."""
    sizes = [(64, 64), (75, 50), (333, 40), (1024, 31), (2, 9), (600, 96), (1500, 16), (4096, 6),
             (130, 3)]
    bads = [torch.from_numpy(_degraded_image(nx, ny, seed=7 + i)[1]).to(DEV)
            for i, (nx, ny) in enumerate(sizes)]
    iters = 3
    single = []
    for b in bads:
        t = b.clone()
        af_ph, rms = ctx.pga(t, iters)
        single.append((t, af_ph.clone(), rms.clone()))
    for rep in range(2):
        imgs = [b.clone() for b in bads]
        res = ctx.pga_scenes(imgs, iters)
        torch.cuda.synchronize()
        for (t, af_ph, rms), img, (a2, r2) in zip(single, imgs, res):
            assert torch.equal(img, t), rep
            assert torch.equal(a2, af_ph), rep
            assert torch.equal(r2, rms), rep



@pytest.mark.parametrize("nx,ny", [(64, 64), (600, 96)])
def test_pga_rms_stop_ends_before_applying(ctx, nx, ny):
    """rms_stop (the C++ restatement's `if (rms < 0.1) break`,
    autofocus_cpu.cpp:432): with the threshold set just above the fixed-count
    run's rms at iteration 3, the run stops at the first iteration k whose
    rms is below it -- rms[0..k] equal the fixed run's, later entries are
    NaN, and the image and af_ph equal a fixed run of k iterations (the
    estimate of iteration k is not applied).  Single and batched entries."""
    _, bad = _degraded_image(nx, ny, seed=11)
    iters = 8
    t = torch.from_numpy(bad.copy()).to(DEV)
    _, rms_fixed = ctx.pga(t, iters)
    rf = rms_fixed.cpu().numpy()
    thr = float(rf[3]) * (1 + 1e-9)
    k = int(np.nonzero(rf < thr)[0][0])
    assert k <= 3
    ref = torch.from_numpy(bad.copy()).to(DEV)
    af_k, rms_k = ctx.pga(ref, k)
    t = torch.from_numpy(bad.copy()).to(DEV)
    af_s, rms_s = ctx.pga(t, iters, rms_stop=thr)
    rs = rms_s.cpu().numpy()
    np.testing.assert_array_equal(rs[:k + 1], rf[:k + 1])
    assert np.all(np.isnan(rs[k + 1:]))
    assert torch.equal(t, ref)
    assert torch.equal(af_s, af_k)
    imgs = [torch.from_numpy(bad.copy()).to(DEV), torch.from_numpy(bad.copy()).to(DEV)]
    res = ctx.pga_scenes(imgs, iters, rms_stop=thr)
    torch.cuda.synchronize()
    for img, (a2, r2) in zip(imgs, res):
        assert torch.equal(img, ref)
        assert torch.equal(a2, af_k)
        np.testing.assert_array_equal(r2.cpu().numpy()[:k + 1], rf[:k + 1])
    # no stop: the fixed-count result, bit for bit
    t2 = torch.from_numpy(bad.copy()).to(DEV)
    _, r3 = ctx.pga(t2, iters, rms_stop=1e-12)
    np.testing.assert_array_equal(r3.cpu().numpy(), rf)
