"""HIP graph capture of a serving batch: range compression of two uneven
scenes (per scene, or batched in one sarf_range_compress_scenes call), their
backprojection in one sarf_bp_scenes launch and PGA (per image, or batched
in one sarf_pga_scenes call), captured
once (after one eager call that uploads the scene list) and replayed
several times with other GPU work and synchronisations in between.  Every
replay must reproduce the eager result exactly (the kernels are
deterministic and the graph holds no memset nodes)."""
import numpy as np
import pytest
import torch

from scene_data import small_inputs
from uncc_sar_focusing_amd.scenes import ImageGrid

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("mode", ["per_scene", "batched"])
def test_serving_batch_graph_replays_match_eager(ctx, mode):
    specs = [(48, 96, ImageGrid(70, 45, 0.3, -0.2, 6.9, 4.4), 1024),
             (31, 64, ImageGrid(40, 52, -0.1, 0.2, 3.9, 5.1), 512)]
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
        if mode == "batched":
            ctx.range_compress_scenes([s["ph"] for s in scenes], [s["nfft"] for s in scenes],
                                      out=[s["prof"] for s in scenes])
        else:
            for s in scenes:
                ctx.range_compress(s["ph"], s["nfft"], out=s["prof"])
        ctx.backproject_scenes(descs)
        if mode == "batched":
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
