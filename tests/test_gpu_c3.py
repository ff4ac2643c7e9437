"""Differentiable autofocus at the full grid-search geometry (512^2 image,
1024 pulses x 1024 samples, Nfft 4096, clutter on every 4th cell) with the
per-pulse 3-D position errors of a smoothed Gaussian random walk (sigma
0.05 m per axis): 3072 unknowns.

* the adjoint backprojection's gradient dE/d(antenna positions) equals the
  FP64 CPU oracle's adjoint (rel-L2 <= 1e-4) and central finite differences
  (h = 1e-4 m) of the all-FP64 GPU forward model on 32 random pulses;
* 20 Adam steps of a smoothness-parametrised descent through the
  torch.autograd harness lower the entropy (overall, and on at least 15 of
  the 19 steps: momentum overshoots now and then), and reduce the
  line-of-sight error once its unobservable constant + linear part (an
  image shift) is removed.
"""
import numpy as np
import pytest
import torch

import oracle
from uncc_sar_focusing_amd import autofocus as af
from uncc_sar_focusing_amd.autograd import BpOperator, gradient_autofocus, los_residual_rms, \
    smoothing_operator

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(scope="module")
def c3(ctx):
    from uncc_sar_focusing_amd.workloads import random_walk_scene

    sc = random_walk_scene(ctx, "C2")
    g, grid = sc.geom, sc.grid
    los = g.los(grid.x0, grid.y0)
    prof = ctx.range_compress(sc.ph, sc.cfg.nfft)
    return sc, prof, los


def test_c3_gradient_matches_cpu_fp64_and_finite_differences(ctx, c3):
    sc, prof, los = c3
    g, grid = sc.geom, sc.grid
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    img = ctx.backproject(prof, sc.ant, sc.r0, sc.f0, grid, g.maxWr, los=los)
    ent = ctx.image_entropy(img)
    grad = ctx.entropy_grad(prof, sc.ant, sc.r0, sc.f0, grid, g.maxWr, img, ent,
                            los=los).cpu().numpy()
    prof64 = ctx.range_compress(sc.ph.to(torch.complex128), sc.cfg.nfft)
    prof_h = prof64.cpu().numpy()
    img_ref = oracle.bp_f64(prof_h, g.ant, g.r0, g.f0, grid, g.maxWr)
    ref = oracle.bp_adjoint_f64(prof_h, g.ant, g.r0, g.f0, grid, g.maxWr, img_ref)
    rel = np.linalg.norm(grad - ref) / np.linalg.norm(ref)
    print(f"C3 adjoint vs CPU FP64 oracle: rel-L2 {rel:.3e}")
    assert rel <= 1e-4

    def E64(a):
        im = ctx.backproject(prof64, t(a), sc.r0, sc.f0, grid, g.maxWr, precision="f64")
        return float(ctx.image_entropy(im).cpu()[0])

    rng = np.random.default_rng(0)
    h = 1e-4
    worst = 0.0
    for n in rng.choice(g.Np, 32, replace=False):
        for ax in range(3):
            ap, am = g.ant.copy(), g.ant.copy()
            ap[n, ax] += h
            am[n, ax] -= h
            fd = (E64(ap) - E64(am)) / (2 * h)
            worst = max(worst, abs(fd - grad[n, ax]))
    gmax = float(np.abs(grad).max())
    print(f"C3 central FD (h = 1e-4 m, 32 pulses): max |FD - adjoint| / max |grad| "
          f"{worst / gmax:.3e}")
    assert worst <= 2e-3 * gmax


def test_c3_adam_steps_lower_entropy_and_los_error(ctx, c3):
    sc, prof, los = c3
    g, grid = sc.geom, sc.grid
    op = BpOperator(ctx, prof, sc.r0, sc.f0, grid, g.maxWr, los)
    smooth = smoothing_operator(g.Np, 16.0, DEV)
    make = lambda q: sc.ant + smooth(q)  # noqa: E731
    q0 = torch.zeros((g.Np, 3), dtype=torch.float64, device=DEV)
    gr = gradient_autofocus(op, make, q0, 20, 0.003)
    e = gr.entropies
    print("C3 entropy over 20 Adam steps:", np.round(e, 5))
    # Adam's steps are not monotone (momentum overshoots now and then), the
    # descent is: the entropy falls over the run and on most steps
    assert e[-1] < e[0] - 0.05, e
    assert np.count_nonzero(np.diff(e) < 0) >= 15, e
    err = sc.true_geom.ant - g.ant
    d = af.los_directions(g.ant, grid.x0, grid.y0)
    before = los_residual_rms(err, d)
    after = los_residual_rms(err - smooth(gr.params).cpu().numpy(), d)
    print(f"C3 detrended LOS error rms: {before:.5f} -> {after:.5f} m")
    assert after < before
