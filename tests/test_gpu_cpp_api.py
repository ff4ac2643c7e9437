"""The C++ drop-in interface end to end on the GPU: the example program
(sph_sar_data_callback_gpu on a Sandia-style chirp record -> BMP, then
cuda_focus_SAR_image with zero padding) must put the brightest pixel on the
strongest synthetic target, and writeBMPFile must produce the reference's
file byte for byte (the oracle's quantisation + the loadbmp layout)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from uncc_sar_focusing_amd.bmp import encode_bmp

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "uncc_sar_focusing_amd", "examples", "example_gpu")


def test_cpp_example_focuses_on_target(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("example not built")
    bmp = tmp_path / "out.bmp"
    raw = tmp_path / "img.bin"
    bmp2 = tmp_path / "img.bmp"
    r = subprocess.run([EXE, str(bmp), str(raw), str(bmp2)], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "peak x=0.000 y=0.000" in r.stdout or "peak x=-0.000" in r.stdout.replace("y=-0", "y=0")
    data = bmp.read_bytes()
    assert data[:2] == b"BM" and len(data) > 54
    # explicit 101 x 101 image: BMP bytes identical to the reference writer
    img = np.fromfile(raw, dtype=np.complex64).reshape(101, 101)
    expect = encode_bmp(oracle.db8(img, 70.0))
    got = bmp2.read_bytes()
    assert len(got) == len(expect)
    assert got == expect


RF = os.path.join(REPO, "uncc_sar_focusing_amd", "examples", "read_focus")


def test_read_focus_sandia_mat_end_to_end(tmp_path):
    """A Sandia-format MAT record (scipy-written, compressed, 2 channels)
    with one point target at (1.0, -2.0) m: readData -> callback -> BMP on
    the GPU, and the default image's brightest pixel lies on the target."""
    import scipy.io

    if not os.path.exists(RF):
        pytest.skip("read_focus not built")
    K, Np, adf = 256, 128, 1.0e6
    f0, bw = 9.6e9, 640e6
    df = bw / (K - 1)
    th = np.deg2rad(90.0 - 2.0 + 4.0 * np.arange(Np) / (Np - 1))
    pos = np.stack([10000.0 * np.cos(th), 10000.0 * np.sin(th), np.full(Np, 7500.0)])
    pos = pos.astype(np.float32).astype(np.float64)  # what the float reader will see
    r0 = np.sqrt((pos ** 2).sum(0))
    tgt = np.array([1.0, -2.0, 0.0])
    dR = np.sqrt(((pos - tgt[:, None]) ** 2).sum(0)) - r0
    fk = f0 + df * np.arange(K)
    ph = np.exp(-4j * np.pi * fk[:, None] * dR[None, :] / 299792458.0)  # K x Np
    sd = np.stack([ph, 0.5 * ph], axis=2)
    rec = {"sph_MATData": {"preamble": {"ADF": np.array([[adf]])},
                           "Data": {"SampleData": sd,
                                    "StartF": np.full((1, Np), f0),
                                    "ChirpRate": np.full((1, Np), df * adf),
                                    "radarCoordinateFrame": {"x": pos[0:1], "y": pos[1:2],
                                                             "z": pos[2:3]}}}}
    mat = tmp_path / "scene.mat"
    scipy.io.savemat(mat, rec, do_compression=True)
    bmp = tmp_path / "scene.bmp"
    r = subprocess.run([RF, str(mat), str(bmp), "100000", "HV"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("image ")][-1]
    vals = dict(kv.split("=") for kv in line.split() if "=" in kv)
    dx, dy = float(vals["dx"]), float(vals["dy"])
    assert abs(float(vals["x"]) - tgt[0]) <= 1.5 * dx
    assert abs(float(vals["y"]) - tgt[1]) <= 1.5 * dy
    data = bmp.read_bytes()
    assert data[:2] == b"BM" and data[28] == 24
