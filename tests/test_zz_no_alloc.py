"""sarf_backproject_noalloc through the C ABI against sarf_backproject on the
same inputs, two calls in a row, the second smaller than the first:

* with the same explicit tile and pulse split both entries run the same
  kernels in the same order, so the images are bit-identical;
* with an automatic shape whose partial images fit the fixed scratch the
  split rule is the same, so again bit-identical;
* where the automatic split count would need more than the 64 MiB scratch,
  the allocation-free entry lowers it and matches sarf_backproject with that
  lower count bit for bit, and stays within 1e-6 rel-L2 of the automatic
  image; an explicit split count that does not fit is rejected.
"""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from uncc_sar_focusing_amd import _lib
from uncc_sar_focusing_amd._lib import BpDesc, SARF_ERR_INVALID
from uncc_sar_focusing_amd.scenes import ImageGrid, circular_aperture, point_targets

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _scene(Np, K, n, spacing, nfft, seed):
    geom = circular_aperture(Np, K)
    tg = point_targets(5, 0.6 * spacing * n, seed=seed)
    ph = oracle.synth_ph_f64(geom, tg).astype(np.complex64)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    return geom, t(ph), t(geom.ant), t(geom.r0), t(geom.f0), ImageGrid.square(n, spacing), nfft


def _call(ctx, fn, prof, ant, r0, f0, grid, maxWr, tile, split, los):
    Np, nfft = prof.shape
    img = torch.full((grid.Nx, grid.Ny), float("nan"), dtype=torch.complex64, device=DEV)
    d = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
               _lib.SARF_BP_MIXED, 0, tile, split, *los)
    rc = fn(ctx._h, C.byref(d), C.c_void_p(prof.data_ptr()), C.c_void_p(ant.data_ptr()),
            C.c_void_p(r0.data_ptr()), C.c_void_p(f0.data_ptr()), C.c_void_p(img.data_ptr()),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return rc, img.cpu().numpy()


@pytest.mark.parametrize("tile,split", [(13, 4), (0, 0)])
def test_noalloc_matches_backproject_two_calls(ctx, tile, split):
    lib = ctx.lib
    for (Np, K, n, sp, nfft, seed) in ((512, 256, 512, 0.05, 2048, 3), (64, 96, 96, 0.1, 512, 4)):
        geom, ph, ant, r0, f0, grid, nfft = _scene(Np, K, n, sp, nfft, seed)
        prof = ctx.range_compress(ph, nfft)
        los = geom.los()
        print(f"noalloc case Np={Np} image {n}^2 tile={tile} split={split}", flush=True)
        rc_a, ref = _call(ctx, lib.sarf_backproject, prof, ant, r0, f0, grid, geom.maxWr, tile,
                          split, los)
        rc_b, got = _call(ctx, lib.sarf_backproject_noalloc, prof, ant, r0, f0, grid, geom.maxWr,
                          tile, split, los)
        assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, lib.sarf_last_error())
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_lowers_automatic_splits_to_fit(ctx):
    """2048^2 image (2048 tiles of tile 13: head + tail splits): the
    automatic 3 groups write 2 partial images of 32 MiB, which fit the 64 MiB
    as they are; 2560^2: 2 partials of 50 MiB do not, so the allocation-free
    entry lowers the count to 2 groups (one partial) and matches
    sarf_backproject with 2 bit for bit."""
    lib = ctx.lib
    geom, ph, ant, r0, f0, grid, nfft = _scene(256, 256, 2048, 0.05, 2048, 5)
    prof = ctx.range_compress(ph, nfft)
    los = geom.los()
    print("noalloc case 2048^2 automatic shape", flush=True)
    rc, got = _call(ctx, lib.sarf_backproject_noalloc, prof, ant, r0, f0, grid, geom.maxWr, 13, 0,
                    los)
    assert rc == 0, lib.sarf_last_error()
    rc, auto = _call(ctx, lib.sarf_backproject, prof, ant, r0, f0, grid, geom.maxWr, 13, 0, los)
    assert rc == 0
    assert np.array_equal(got.view(np.uint32), auto.view(np.uint32))
    rc, three = _call(ctx, lib.sarf_backproject, prof, ant, r0, f0, grid, geom.maxWr, 13, 3, los)
    assert rc == 0
    assert np.array_equal(got.view(np.uint32), three.view(np.uint32))
    print("noalloc case 2560^2 automatic shape (lowered)", flush=True)
    geom5, ph5, ant5, r05, f05, grid5, _ = _scene(256, 256, 2560, 0.05, 2048, 7)
    prof5 = ctx.range_compress(ph5, nfft)
    los5 = geom5.los()
    rc, low = _call(ctx, lib.sarf_backproject_noalloc, prof5, ant5, r05, f05, grid5, geom5.maxWr,
                    13, 0, los5)
    assert rc == 0, lib.sarf_last_error()
    rc, two = _call(ctx, lib.sarf_backproject, prof5, ant5, r05, f05, grid5, geom5.maxWr, 13, 2,
                    los5)
    assert rc == 0
    assert np.array_equal(low.view(np.uint32), two.view(np.uint32))
    rc, auto5 = _call(ctx, lib.sarf_backproject, prof5, ant5, r05, f05, grid5, geom5.maxWr, 13, 0,
                      los5)
    assert rc == 0
    assert oracle.rel_l2(low, auto5) < 1e-6
    del prof5, ph5
    # an explicit split count whose partial images exceed the scratch
    rc, _ = _call(ctx, lib.sarf_backproject_noalloc, prof, ant, r0, f0, grid, geom.maxWr, 13, 8,
                  los)
    assert rc == SARF_ERR_INVALID
    assert b"more than the" in lib.sarf_last_error()
    # a smaller call afterwards still matches
    geom, ph, ant, r0, f0, grid, nfft = _scene(48, 64, 40, 0.1, 512, 6)
    prof = ctx.range_compress(ph, nfft)
    rc_a, ref = _call(ctx, lib.sarf_backproject, prof, ant, r0, f0, grid, geom.maxWr, 0, 0,
                      geom.los())
    rc_b, got = _call(ctx, lib.sarf_backproject_noalloc, prof, ant, r0, f0, grid, geom.maxWr, 0, 0,
                      geom.los())
    assert rc_a == 0 and rc_b == 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
