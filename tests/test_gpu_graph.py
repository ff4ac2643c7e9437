"""HIP graph capture of a serving batch: range compression of two uneven
scenes (per scene, or batched in one sarf_range_compress_scenes call), their
backprojection in one sarf_bp_scenes launch (one tile for both, or a tile
per scene) and PGA (per image, or batched in one sarf_pga_scenes call),
captured
once (after one eager call that uploads the scene list) and replayed
several times with other GPU work and synchronisations in between.  Every
replay must reproduce the eager result exactly (the kernels are
deterministic and the graph holds no memset nodes)."""
import numpy as np
import pytest
import torch

from scene_data import small_inputs
from uncc_sar_focusing_amd import _lib
from uncc_sar_focusing_amd.scenes import ImageGrid

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("mode", ["per_scene", "batched", "batched_mixed"])
def test_serving_batch_graph_replays_match_eager(ctx, mode):
    if mode == "batched_mixed":
        # a tile per scene (fine spacings: tiles 15, 16, 13 are accurate)
        specs = [(48, 96, ImageGrid(270, 45, 0.3, -0.2, 269 * 0.02, 44 * 0.02), 1024),
                 (31, 64, ImageGrid(140, 52, -0.1, 0.2, 139 * 0.02, 51 * 0.02), 512)]
    else:
        specs = [(48, 96, ImageGrid(70, 45, 0.3, -0.2, 6.9, 4.4), 1024),
                 (31, 64, ImageGrid(40, 52, -0.1, 0.2, 3.9, 5.1), 512)]
    tile = _lib.SARF_SCENES_TILE_MIXED if mode == "batched_mixed" else 0
    scenes = []
    for Np, K, grid, nfft in specs:
        geom, tg, ph = small_inputs(Np=Np, K=K)
        scenes.append(dict(grid=grid, nfft=nfft, maxWr=geom.maxWr, ph=_dev(ph),
                           ant=_dev(geom.ant), r0=_dev(geom.r0), f0=_dev(geom.f0),
                           prof=torch.empty((Np, nfft), dtype=torch.complex64, device=DEV),
                           img=torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64,
                                           device=DEV)))
    descs = [(s["prof"], s["ant"], s["r0"], s["f0"], s["grid"], s["maxWr"], s["img"])
             for s in scenes]

    outs = [(torch.empty(s["grid"].Nx, dtype=torch.float64, device=DEV),
             torch.empty(2, dtype=torch.float64, device=DEV)) for s in scenes]

    def batch():
        if mode != "per_scene":
            ctx.range_compress_scenes([s["ph"] for s in scenes], [s["nfft"] for s in scenes],
                                      out=[s["prof"] for s in scenes])
        else:
            for s in scenes:
                ctx.range_compress(s["ph"], s["nfft"], out=s["prof"])
        ctx.backproject_scenes(descs, tile=tile)
        if mode != "per_scene":
            ctx.pga_scenes([s["img"] for s in scenes], 2, out=outs)
        else:
            for s in scenes:
                ctx.pga(s["img"], 2)

    batch()  # eager: uploads the scene list, sizes every scratch buffer
    torch.cuda.synchronize()
    ref = [s["img"].clone() for s in scenes]
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        batch()  # warm the capture stream as well
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        batch()
    other = scenes[0]
    for rep in range(3):
        for s in scenes:
            s["img"].zero_()
        g.replay()
        torch.cuda.synchronize()
        for s, r in zip(scenes, ref):
            assert torch.equal(s["img"], r), rep
        # unrelated eager work between replays
        ctx.backproject(other["prof"], other["ant"], other["r0"], other["f0"], other["grid"],
                        other["maxWr"])
        torch.cuda.synchronize()


def _batched_scenes(specs):
    scenes = []
    for Np, K, grid, nfft in specs:
        geom, tg, ph = small_inputs(Np=Np, K=K)
        scenes.append(dict(grid=grid, nfft=nfft, maxWr=geom.maxWr, ph=_dev(ph),
                           ant=_dev(geom.ant), r0=_dev(geom.r0), f0=_dev(geom.f0),
                           prof=torch.empty((Np, nfft), dtype=torch.complex64, device=DEV),
                           img=torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64,
                                           device=DEV),
                           out=(torch.empty(grid.Nx, dtype=torch.float64, device=DEV),
                                torch.empty(2, dtype=torch.float64, device=DEV))))
    return scenes


def _run_batched(ctx, scenes):
    ctx.range_compress_scenes([s["ph"] for s in scenes], [s["nfft"] for s in scenes],
                              out=[s["prof"] for s in scenes])
    ctx.backproject_scenes([(s["prof"], s["ant"], s["r0"], s["f0"], s["grid"], s["maxWr"],
                             s["img"]) for s in scenes])
    ctx.pga_scenes([s["img"] for s in scenes], 2, out=[s["out"] for s in scenes])


def test_graph_survives_calls_with_other_and_larger_lists(ctx):
    """A graph captured with one scene list keeps replaying it exactly after
    eager batched calls with another, larger list (each list's device state
    is its own; nothing the graph points to is overwritten or freed), and
    after sarf_ctx_trim the same eager call uploads again and still agrees."""
    a = _batched_scenes([(40, 64, ImageGrid(48, 40, 0.2, -0.1, 4.7, 3.9), 512),
                         (33, 96, ImageGrid(36, 52, -0.1, 0.3, 3.5, 5.1), 1024)])
    b = _batched_scenes([(64, 128, ImageGrid(150, 120, 0.0, 0.0, 14.9, 11.9), 2048),
                         (48, 96, ImageGrid(96, 140, 0.1, 0.1, 9.5, 13.9), 1024),
                         (37, 64, ImageGrid(64, 64, 0.0, 0.0, 6.3, 6.3), 512)])
    _run_batched(ctx, a)
    torch.cuda.synchronize()
    ref_a = [s["img"].clone() for s in a]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        _run_batched(ctx, a)
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _run_batched(ctx, a)
    _run_batched(ctx, b)  # another, larger list: uploads its own state
    torch.cuda.synchronize()
    ref_b = [s["img"].clone() for s in b]
    for rep in range(3):
        for s in a:
            s["img"].zero_()
        g.replay()
        torch.cuda.synchronize()
        for s, r in zip(a, ref_a):
            assert torch.equal(s["img"], r), rep
        _run_batched(ctx, b)
        torch.cuda.synchronize()
        for s, r in zip(b, ref_b):
            assert torch.equal(s["img"], r), rep
    del g
    torch.cuda.synchronize()
    ctx.lib.sarf_ctx_trim(ctx._h)
    _run_batched(ctx, a)
    torch.cuda.synchronize()
    for s, r in zip(a, ref_a):
        assert torch.equal(s["img"], r)


def test_one_context_on_two_streams_is_ordered(ctx):
    """Calls of one context on two streams with no synchronisation between
    them share the context's partial-image scratch: the second stream waits
    for the first call, so both images equal single-stream results."""
    geom1, _, ph1 = small_inputs(Np=256, K=128)
    geom2, _, ph2 = small_inputs(Np=192, K=96)
    grid = ImageGrid(128, 96, 0.0, 0.0, 12.7, 9.5)
    p1 = ctx.range_compress(_dev(ph1), 1024)
    p2 = ctx.range_compress(_dev(ph2), 1024)
    args1 = (p1, _dev(geom1.ant), _dev(geom1.r0), _dev(geom1.f0), grid, geom1.maxWr)
    args2 = (p2, _dev(geom2.ant), _dev(geom2.r0), _dev(geom2.f0), grid, geom2.maxWr)
    r1 = ctx.backproject(*args1, pulse_split=8)
    r2 = ctx.backproject(*args2, pulse_split=8)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(torch.cuda.current_stream())
    s2.wait_stream(torch.cuda.current_stream())
    for rep in range(3):
        with torch.cuda.stream(s1):
            o1 = ctx.backproject(*args1, pulse_split=8)
        with torch.cuda.stream(s2):
            o2 = ctx.backproject(*args2, pulse_split=8)
        torch.cuda.synchronize()
        assert torch.equal(o1, r1) and torch.equal(o2, r2), rep
