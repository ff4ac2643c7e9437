"""The multi-resolution grid-search driver (sarf_grid_search) on the GPU.

* Every level's box follows the reference's shrink rule
  (src/autofocus/autofocus_gpu.cu:938-939: start' = m - w/4, then the end
  from the already-moved start) or the centred one; every level's minimiser
  and entropy equal an independent sarf_bp_batch_entropy evaluation of that
  level's box (same kernel: bit-identical), and the reported covariance is
  the host estimator applied to those values.
* The reference's own trajectory models (D = 7 linear, D = 10 quadratic in
  the pulse index, gridSearchErrorFunctions_dft.cuh:279-291) around the
  bestFit / quadFit centre of a recorded track (autofocus_gpu.cu:842-881,
  1000-1029): the same level-by-level agreement with the batched entropy,
  and the minimiser's entropy equals the FP64 oracle's (1e-4 absolute).
* The detrended-basis search on a clutter-free scene lands within one
  final-level step of the injected line-of-sight error.
"""
import numpy as np
import pytest
import torch

import oracle
from uncc_sar_focusing_amd import autofocus as af
from uncc_sar_focusing_amd.scenes import Geometry, ImageGrid, circular_aperture, point_targets

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _box_points(start, end, num):
    """Candidates in the driver's order (first axis slowest), each axis point
    start + i ((end - start) / (num - 1)) as the driver forms it."""
    axes = []
    for s, e, n in zip(start, end, num):
        axes.append(np.array([s]) if n == 1 else s + np.arange(n) * ((e - s) / (n - 1)))
    return np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, len(num))


TRUE_COEF = (0.01, -0.015, 0.012)


def _los_scene(Np=64, K=96, n=40, spacing=0.1, true_coef=TRUE_COEF, seed=5):
    geom = circular_aperture(Np, K)
    tg = point_targets(4, n * spacing * 0.5, seed=seed)
    basis = af.detrended_basis(af.poly_basis(Np))
    d = af.los_directions(geom.ant)
    true_ant = geom.ant + (np.asarray(true_coef) @ basis)[:, None] * d
    true_geom = Geometry(ant=true_ant, r0=geom.r0, f0=geom.f0, df=geom.df, K=geom.K)
    ph = oracle.synth_ph_f64(true_geom, tg).astype(np.complex64)
    return geom, ph, ImageGrid.square(n, spacing), basis, d


@pytest.mark.parametrize("shrink", ["reference", "centred"])
def test_levels_follow_shrink_rule_and_match_batch_entropy(ctx, shrink):
    geom, ph, grid, basis, d = _los_scene()
    nfft = 1024
    prof = ctx.range_compress(to_dev(ph), nfft)
    ant, r0, f0 = to_dev(geom.ant), to_dev(geom.r0), to_dev(geom.f0)
    tb, td = to_dev(basis), to_dev(d)
    start, end, num = np.array([-0.03, -0.03, -0.03]), np.array([0.03, 0.03, 0.03]), [5, 4, 7]
    res = af.grid_search_multires(ctx, prof, r0, f0, grid, geom.maxWr, start, end, num, 4,
                                  ant=ant, basis=tb, direction=td, shrink=shrink)
    assert len(res.levels) == 4
    st, en = start.copy(), end.copy()
    for lv, L in enumerate(res.levels):
        np.testing.assert_array_equal(L["start"], st)
        np.testing.assert_array_equal(L["end"], en)
        P = _box_points(st, en, num)
        assert L["ncand"] == len(P)
        e = ctx.batch_entropy(prof, ant, r0, f0, grid, geom.maxWr, to_dev(P), tb,
                              td).cpu().numpy()
        k = int(np.argmin(e))
        np.testing.assert_array_equal(L["best"], P[k])
        assert L["entropy"] == e[k]
        assert L["gpu_ms"] > 0
        if lv == len(res.levels) - 1:
            cov = af.grid_covariance(st, en, num, e, k)
            np.testing.assert_array_equal(np.isnan(res.covariance), np.isnan(cov))
            np.testing.assert_allclose(res.covariance, cov, rtol=1e-12, atol=0)
        m = P[k]
        w = en - st
        st = m - w / 4
        en = m + (en - st) / 4 if shrink == "reference" else m + w / 4
    np.testing.assert_array_equal(res.best, res.levels[-1]["best"])
    assert res.best_entropy == res.levels[-1]["entropy"]
    # clutter-free scene, error inside the first box: the search lands within
    # one final-level step of the injected error (the FP64 oracle's entropy
    # over the same boxes does too; with a 0.2 m box around an error of
    # 0.05-0.08 m the coarse first level picks another basin, in both)
    step = (res.levels[-1]["end"] - res.levels[-1]["start"]) / (np.array(num) - 1)
    assert np.all(np.abs(res.best - np.array(TRUE_COEF)) <= step + 1e-12), (res.best, step)


def _track(Np, quadratic):
    i = np.arange(Np, dtype=np.float64)
    x = -352.0 + 11.0 * i
    y = np.full(Np, 10000.0) + (0.03125 * (i - Np / 2) ** 2 if quadratic else 0.0)
    z = np.full(Np, 7500.0)
    return np.stack([x, y, z], axis=1)


@pytest.mark.parametrize("D", [7, 10])
def test_reference_trajectory_model_levels(ctx, D):
    """The reference's D = 7 / D = 10 parameter vectors through the driver:
    the first box is centred on the bestFit / quadFit fit of a recorded
    track that is exactly linear / quadratic in the pulse index (its centre
    reproduces the track), spans narrowed and fewer samples than the
    reference's 23^3 x 5 / 5^6 x 5 so the test stays small; every level's
    minimiser and entropy equal sarf_bp_batch_entropy over the same box
    through the Python mapping of the parameters
    (reference_trajectory_candidates), and the minimiser's entropy equals the
    FP64 oracle's image entropy at those positions (1e-4).  (Which candidate
    wins is the entropy's business: on a small scene a velocity error that
    shifts targets out of the image can beat the recorded track.)"""
    Np, K, nfft = 64, 96, 1024
    ant = _track(Np, D == 10)
    r0 = np.linalg.norm(ant, axis=1)
    base = circular_aperture(Np, K)
    geom = Geometry(ant=ant, r0=r0, f0=base.f0, df=base.df, K=K)
    grid = ImageGrid.square(32, 0.1)
    tg = point_targets(3, 2.0, seed=11)
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    prof_h = oracle.range_compress_f64(ph, nfft)
    prof = to_dev(prof_h.astype(np.complex64))
    r0d, f0d = to_dev(r0), to_dev(geom.f0)
    start, end, num = af.reference_grid_box(ant, D)
    centre = (start + end) / 2
    np.testing.assert_allclose(oracle.reference_trajectory_f64(centre, D, Np), ant, rtol=0,
                               atol=2e-3)
    half = np.where(num > 1, (end - start) / 2, 0.0) * (0.004 if D == 7 else 0.002)
    start, end = centre - half, centre + half
    num = np.where(num > 1, 3, 1)
    res = af.grid_search_multires(ctx, prof, r0d, f0d, grid, geom.maxWr, start, end, num, 2,
                                  model="ref7" if D == 7 else "ref10")
    zero = to_dev(np.zeros_like(ant))
    for L in res.levels:
        P = _box_points(L["start"], L["end"], num)
        assert L["ncand"] == len(P)
        basis, direction, coefs = af.reference_trajectory_candidates(P, D, Np)
        e = ctx.batch_entropy(prof, zero, r0d, f0d, grid, geom.maxWr, to_dev(coefs),
                              to_dev(basis), to_dev(direction)).cpu().numpy()
        k = int(np.argmin(e))
        np.testing.assert_array_equal(L["best"], P[k])
        assert L["entropy"] == e[k]
    pos = oracle.reference_trajectory_f64(res.best, D, Np)
    ref = oracle.entropy_f64(oracle.bp_f64(prof_h, pos, r0, geom.f0, grid, geom.maxWr))
    assert abs(res.best_entropy - ref) < 1e-4


def test_grid_search_rejects_bad_boxes(ctx):
    geom, ph, grid, basis, d = _los_scene(Np=32, K=64, n=16)
    prof = ctx.range_compress(to_dev(ph), 512)
    args = (prof, to_dev(geom.r0), to_dev(geom.f0), grid, geom.maxWr)
    kw = dict(ant=to_dev(geom.ant), basis=to_dev(basis), direction=to_dev(d))
    from uncc_sar_focusing_amd._lib import SarfError

    with pytest.raises(SarfError):  # end < start
        af.grid_search_multires(ctx, *args, [0.1, 0, 0], [0.0, 1, 1], [3, 3, 3], 1, **kw)
    with pytest.raises(SarfError):  # too many candidates per level
        af.grid_search_multires(ctx, *args, [0, 0, 0], [1, 1, 1], [4096, 4096, 2], 1, **kw)
    with pytest.raises(SarfError):  # reference model with the wrong parameter count
        af.grid_search_multires(ctx, *args, [0] * 3, [1] * 3, [2] * 3, 1, model="ref7")
