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
