"""GPU parity of the entropy reduction and the batched grid-search objective
against the FP64 oracle (entropy of the FP64 oracle image of each candidate
trajectory).  Tolerances: image entropy 1e-9 (FP64 path) / 1e-6 (FP32
image); candidate entropies within 1e-4 absolute; identical argmin."""
import numpy as np
import pytest
import torch

import oracle
from scene_data import small_inputs
from uncc_sar_focusing_amd import autofocus as af
from uncc_sar_focusing_amd.scenes import ImageGrid, Targets, circular_aperture, point_targets

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_image_entropy_matches_oracle(ctx, dtype):
    rng = np.random.default_rng(2)
    img = (rng.standard_normal((300, 217)) + 1j * rng.standard_normal((300, 217))).astype(dtype)
    img[5:20] *= 30.0
    out = ctx.image_entropy(to_dev(img)).cpu().numpy()
    ref = oracle.entropy_f64(img)
    tol = 1e-12 if dtype == np.complex128 else 1e-9
    assert abs(out[0] - ref) < tol * 10
    m = np.abs(img.astype(np.complex128)) ** 2
    assert abs(out[1] - m.sum()) < 1e-9 * m.sum()


def _los_scene(Np=48, K=64, n=40, spacing=0.12, true_coef=(0.02, -0.015, 0.01)):
    geom = circular_aperture(Np, K)
    tg = point_targets(5, n * spacing * 0.6, seed=21)
    true_geom = af.perturbed(geom, true_coef)
    ph = oracle.synth_ph_f64(true_geom, tg).astype(np.complex64)
    grid = ImageGrid.square(n, spacing)
    return geom, ph, grid


def test_batch_entropy_reference_trajectory_model(ctx):
    """The reference's own candidate model (absolute positions quadratic in
    the pulse index, D = 10, gridSearchErrorFunctions_dft.cuh:279-291) through
    the per-basis form: each candidate's batched entropy equals the
    single-image path's at those positions (1e-5 relative) and the FP64
    oracle's (1e-4 absolute), with the same argmin."""
    geom, ph, grid = _los_scene()
    nfft = 512
    i = np.arange(geom.Np, dtype=np.float64)
    base = np.zeros(10)
    for a in range(3):  # least-squares quadratic in i per axis: (c0, c1, c2)
        c2, c1, c0 = np.polyfit(i, geom.ant[:, a], 2)
        base[3 * a:3 * a + 3] = (c0, c1, c2)
    base[9] = 1.0
    params = np.repeat(base[None, :], 7, axis=0)
    params[:, 5] += np.linspace(-2e-5, 2e-5, 7)   # y curvature
    params[1::2, 9] *= 1.0 + 1e-5                 # velocity scale
    basis, direction, coefs = af.reference_trajectory_candidates(params, 10, geom.Np)
    prof_h = oracle.range_compress_f64(ph, nfft)
    prof = to_dev(prof_h.astype(np.complex64))
    r0, f0 = to_dev(geom.r0), to_dev(geom.f0)
    ent = ctx.batch_entropy(prof, to_dev(np.zeros_like(geom.ant)), r0, f0, grid, geom.maxWr,
                            to_dev(coefs), to_dev(basis), to_dev(direction)).cpu().numpy()
    ref = []
    for j in range(len(params)):
        pos = oracle.reference_trajectory_f64(params[j], 10, geom.Np)
        img = ctx.backproject(prof, to_dev(pos), r0, f0, grid, geom.maxWr)
        e = float(ctx.image_entropy(img).cpu()[0])
        assert abs(ent[j] - e) <= 1e-5 * abs(e), (j, ent[j], e)
        ref.append(oracle.entropy_f64(oracle.bp_f64(prof_h, pos, geom.r0, geom.f0, grid,
                                                    geom.maxWr)))
    assert np.max(np.abs(ent - np.array(ref))) < 1e-4
    assert int(np.argmin(ent)) == int(np.argmin(ref))


@pytest.mark.parametrize("tile", [0, 4, 11, 12, 13, 14, 15, 16])
def test_batch_entropy_matches_oracle_per_candidate(ctx, tile):
    """Every batch tile shape, both lane orientations (0 = default choice)."""
    geom, ph, grid = _los_scene()
    nfft = 512
    basis = af.poly_basis(geom.Np)
    direction = af.los_directions(geom.ant)
    coefs = af.coef_grid(-0.03, 0.03, 3)  # 27 candidates, includes the truth's neighbourhood
    prof_h = oracle.range_compress_f64(ph, nfft)
    prof = to_dev(prof_h.astype(np.complex64))
    ent = ctx.batch_entropy(prof, to_dev(geom.ant), to_dev(geom.r0), to_dev(geom.f0), grid,
                            geom.maxWr, to_dev(coefs), to_dev(basis), to_dev(direction),
                            tile=tile).cpu().numpy()
    ref = []
    for c in coefs:
        ant = af.apply_los_offsets(geom.ant, c, basis, direction)
        img = oracle.bp_f64(prof_h, ant, geom.r0, geom.f0, grid, geom.maxWr)
        ref.append(oracle.entropy_f64(img))
    ref = np.array(ref)
    assert np.max(np.abs(ent - ref)) < 1e-4, np.max(np.abs(ent - ref))
    assert int(np.argmin(ent)) == int(np.argmin(ref))


def test_batch_zero_candidate_equals_backprojected_image_entropy(ctx):
    geom, tg, ph = small_inputs(Np=64, K=96)
    grid = ImageGrid(70, 45, 0.4, -0.3, 7.0, 5.5)  # ragged tiles
    prof = ctx.range_compress(to_dev(ph), 1024)
    ant, r0, f0 = to_dev(geom.ant), to_dev(geom.r0), to_dev(geom.f0)
    basis = to_dev(af.poly_basis(geom.Np))
    direction = to_dev(af.los_directions(geom.ant))
    coefs = to_dev(np.zeros((3, 3)))
    ent = ctx.batch_entropy(prof, ant, r0, f0, grid, geom.maxWr, coefs, basis, direction).cpu().numpy()
    img = ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr)
    e_img = ctx.image_entropy(img).cpu().numpy()[0]
    assert np.all(ent == ent[0])  # deterministic across candidates
    assert abs(ent[0] - e_img) < 1e-6


def test_grid_search_finds_injected_los_error(ctx):
    """Tiny grid (3^3, 32 pulses, 64^2): the argmin is the injected error and
    matches the FP64 brute force."""
    true = (0.02, 0.0, -0.02)
    geom, ph, grid = _los_scene(Np=32, K=64, n=64, spacing=0.1, true_coef=true)
    nfft = 512
    prof = ctx.range_compress(to_dev(ph), nfft)
    basis = af.poly_basis(geom.Np)
    direction = af.los_directions(geom.ant)
    coefs = af.coef_grid(-0.02, 0.02, 3)
    res = af.grid_search(ctx, prof, to_dev(geom.ant), to_dev(geom.r0), to_dev(geom.f0), grid,
                         geom.maxWr, to_dev(basis), to_dev(direction), coefs)
    assert np.allclose(res.best_coef, true)
    prof_h = oracle.range_compress_f64(ph, nfft)
    ref = [oracle.entropy_f64(oracle.bp_f64(prof_h, af.apply_los_offsets(geom.ant, c, basis,
                                                                          direction),
                                            geom.r0, geom.f0, grid, geom.maxWr)) for c in coefs]
    assert int(np.argmin(ref)) == int(np.argmin(res.entropies))


# ---------------------------------------------------------------- adjoint
def _misfocused(Np=40, K=64, n=32, spacing=0.15, seed=31, f0=9.6e9):
    geom = circular_aperture(Np, K, f0=f0)
    tg = point_targets(6, n * spacing * 0.7, seed=seed)
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    ant = geom.ant.copy()
    t = np.linspace(-1, 1, Np)
    ant[:, 0] += 0.004 * t ** 2
    ant[:, 1] += 0.006 * t ** 3 - 0.003 * t
    ant[:, 2] += 0.002 * t ** 2
    return geom, ph, ant, ImageGrid.square(n, spacing)


@pytest.mark.parametrize("f0,tile,spacing", [(9.6e9, 0, 0.15), (35e9, 0, 0.15), (9.6e9, 4, 0.15),
                                             (9.6e9, 13, 0.15), (9.6e9, 11, 0.15),
                                             (35e9, 13, 0.15), (9.6e9, 16, 0.15),
                                             (9.6e9, 15, 0.04), (35e9, 15, 0.04)])
def test_adjoint_matches_fp64_oracle_gradient(ctx, f0, tile, spacing):
    """GPU adjoint (mixed precision) vs the FP64 CPU adjoint of the FP64 image,
    at X band and at a Ka-band carrier whose per-tile phase span exceeds the
    sine unit's input range.  Tolerance: 2e-3 relative in the 2-norm over all
    pulses and axes.  (Tile 15, 256 pixels wide, needs the finer spacing for
    its profile window to fit in LDS; at 0.15 m it is rejected.)"""
    geom, ph, ant, grid = _misfocused(f0=f0, spacing=spacing)
    nfft = 1024
    prof_h = oracle.range_compress_f64(ph, nfft)
    prof = to_dev(prof_h.astype(np.complex64))
    r0, f0 = to_dev(geom.r0), to_dev(geom.f0)
    img = ctx.backproject(prof, to_dev(ant), r0, f0, grid, geom.maxWr)
    ent = ctx.image_entropy(img)
    g = ctx.entropy_grad(prof, to_dev(ant), r0, f0, grid, geom.maxWr, img, ent,
                         tile=tile).cpu().numpy()
    img_ref = oracle.bp_f64(prof_h, ant, geom.r0, geom.f0, grid, geom.maxWr)
    g_ref = oracle.bp_adjoint_f64(prof_h, ant, geom.r0, geom.f0, grid, geom.maxWr, img_ref)
    rel = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    assert rel < 2e-3, rel


def test_adjoint_matches_central_differences_on_32_pulses(ctx):
    """Central differences (h = 1e-4 m) of the FP64 GPU forward model on 32
    random pulses x 3 axes vs the adjoint.  Tolerance 2e-3 of max |grad|."""
    geom, ph, ant, grid = _misfocused(Np=96, K=96, n=40, spacing=0.12)
    nfft = 1024
    prof_h = oracle.range_compress_f64(ph, nfft)
    prof32 = to_dev(prof_h.astype(np.complex64))
    prof64 = to_dev(prof_h)
    r0, f0 = to_dev(geom.r0), to_dev(geom.f0)
    img = ctx.backproject(prof32, to_dev(ant), r0, f0, grid, geom.maxWr)
    g = ctx.entropy_grad(prof32, to_dev(ant), r0, f0, grid, geom.maxWr, img,
                         ctx.image_entropy(img)).cpu().numpy()

    def E(a):
        im = ctx.backproject(prof64, to_dev(a), r0, f0, grid, geom.maxWr, precision="f64")
        return float(ctx.image_entropy(im).cpu()[0])

    rng = np.random.default_rng(7)
    pulses = rng.choice(geom.Np, 32, replace=False)
    h = 1e-4
    gmax = np.abs(g).max()
    for n in pulses:
        for ax in range(3):
            ap, am = ant.copy(), ant.copy()
            ap[n, ax] += h
            am[n, ax] -= h
            fd = (E(ap) - E(am)) / (2 * h)
            assert abs(fd - g[n, ax]) <= 2e-3 * gmax, (n, ax, fd, g[n, ax], gmax)


def test_autograd_chain_rule_through_los_polynomial(ctx):
    """torch autograd through ant(c) = ant0 + (c . basis) dir equals the
    explicit chain rule of the per-pulse gradient."""
    from uncc_sar_focusing_amd.autograd import BpOperator, entropy

    geom, ph, ant, grid = _misfocused()
    prof = ctx.range_compress(to_dev(ph), 1024)
    op = BpOperator(ctx, prof, to_dev(geom.r0), to_dev(geom.f0), grid, geom.maxWr)
    basis = to_dev(af.poly_basis(geom.Np))
    direction = to_dev(af.los_directions(geom.ant))
    ant0 = to_dev(ant)
    c = torch.tensor([0.01, -0.02, 0.005], dtype=torch.float64, device=DEV, requires_grad=True)
    e = entropy(ant0 + (c @ basis)[:, None] * direction, op)
    e.backward()
    a = (ant0 + (c.detach() @ basis)[:, None] * direction).contiguous()
    img = op.image(a)
    g = ctx.entropy_grad(prof, a, op.r0, op.f0, grid, geom.maxWr, img, ctx.image_entropy(img))
    expect = basis @ (g * direction).sum(dim=1)
    assert torch.allclose(c.grad, expect, rtol=1e-9, atol=1e-12)
