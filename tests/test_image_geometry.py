"""Default image geometry of the C++ interface headers, integer for integer.

The reference sets the width of its default image as
    N_x_pix = (int)((float)Wx_m * N_y_pix) / Wy_m
(include/uncc_sar_focusing.hpp:461, include/charlotte_sar_api.hpp:189): the
float product is truncated before the division (in the precision of the
parameter type) and the quotient truncated again on assignment.  The
expected values below restate that expression with numpy's IEEE float32 /
float64 arithmetic and are compared with what the headers compute, compiled
by g++, on 10^5 seeded geometries and on a known case where dividing first
gives another width.  CPU only.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("nxpix") / "nxpix_check")
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{os.path.join(REPO, 'include')}",
                    os.path.join(HERE, "host", "nxpix_check.cpp"), "-o", out], check=True)
    return out


def ref_width_f32(wx, ny, wy):
    """(int)((float)Wx * Ny) / Wy with float Wx, Wy (all float32 arithmetic)."""
    wx = np.asarray(wx, np.float32)
    wy = np.asarray(wy, np.float32)
    prod = wx * np.asarray(ny).astype(np.float32)
    assert prod.dtype == np.float32
    num = np.trunc(prod).astype(np.int64)
    q = num.astype(np.float32) / wy
    assert q.dtype == np.float32
    return np.trunc(q).astype(np.int64)


def ref_width_f64(wx, ny, wy):
    """The same expression with double Wx, Wy: float product, double quotient."""
    wx32 = np.asarray(wx, np.float64).astype(np.float32)
    prod = wx32 * np.asarray(ny).astype(np.float32)
    num = np.trunc(prod).astype(np.int64)
    return np.trunc(num.astype(np.float64) / np.asarray(wy, np.float64)).astype(np.int64)


def _run(exe, mode, rec: bytes, tmp_path):
    fin, fout = tmp_path / f"{mode}.in", tmp_path / f"{mode}.out"
    fin.write_bytes(rec)
    r = subprocess.run([exe, mode, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return fout.read_bytes()


def test_known_case_truncates_before_dividing(exe, tmp_path):
    """Wx = 78.488 m, N_y = 989, Wy = 236.659 m: the reference gives 327;
    dividing before truncating would give 328."""
    rec = np.zeros(1, dtype=[("wx", "<f4"), ("ny", "<i4"), ("wy", "<f4")])
    rec["wx"], rec["ny"], rec["wy"] = 78.488, 989, 236.659
    got = np.frombuffer(_run(exe, "helper", rec.tobytes(), tmp_path), dtype="<i4")
    assert ref_width_f32(78.488, 989, 236.659) == 327
    assert int(np.float32(78.488) * np.float32(989) / np.float32(236.659)) == 328
    assert list(got) == [327, 327]


def test_helper_matches_reference_expression_on_1e5_geometries(exe, tmp_path):
    rng = np.random.default_rng(20240611)
    n = 100_000
    rec = np.zeros(n, dtype=[("wx", "<f4"), ("ny", "<i4"), ("wy", "<f4")])
    rec["ny"] = rng.integers(16, 8192, n)
    rec["wx"] = rng.uniform(1.0, 3000.0, n)
    rec["wy"] = rng.uniform(1.0, 3000.0, n)
    got = np.frombuffer(_run(exe, "helper", rec.tobytes(), tmp_path), dtype="<i4").reshape(n, 2)
    exp32 = ref_width_f32(rec["wx"], rec["ny"], rec["wy"])
    exp64 = ref_width_f64(rec["wx"].astype(np.float64), rec["ny"], rec["wy"].astype(np.float64))
    assert np.array_equal(got[:, 0], exp32)
    assert np.array_equal(got[:, 1], exp64)
    # the test discriminates: dividing first differs on a share of these
    div_first = np.trunc(rec["wx"] * rec["ny"].astype(np.float32) / rec["wy"]).astype(np.int64)
    assert np.count_nonzero(div_first != exp32) > 50


def test_default_image_of_both_headers_on_1e5_apertures(exe, tmp_path):
    """create_SARImageFormationParams (charlotte_sar_api.hpp) and
    SAR_ImageFormationParameters<T>::create (uncc_sar_focusing.hpp) on
    apertures of realistic statistics: N_fft = N_y = K, Wy = c / (2 mean
    deltaF), Wx = c / (2 |mean dAz| mean startF), N_x by the reference
    expression from the Wx, Wy they hold."""
    rng = np.random.default_rng(7)
    n = 100_000
    K = rng.integers(64, 4097, n).astype(np.float32)
    bw = rng.uniform(100e6, 2e9, n)
    rec = np.stack([K, (bw / (K - 1)).astype(np.float32),
                    rng.uniform(2e-6, 2e-3, n).astype(np.float32),
                    rng.uniform(1e9, 40e9, n).astype(np.float32)], axis=1).astype("<f4")
    raw = _run(exe, "aperture", rec.tobytes(), tmp_path)
    dt = np.dtype([("wx_a", "<f4"), ("wy_a", "<f4"), ("nx_a", "<i4"),
                   ("wx_b", "<f4"), ("wy_b", "<f4"), ("nx_b", "<i4"),
                   ("wx_c", "<f8"), ("wy_c", "<f8"), ("nx_c", "<i4")])
    got = np.frombuffer(raw, dtype=dt)
    assert got.shape[0] == n
    c = 299792458.0
    ny = K.astype(np.int64)
    # extents: c/(2 deltaF) and c/(2 |dAz| startF) in double, stored in T
    wy = c / (2.0 * rec[:, 1].astype(np.float64))
    wx = c / (2.0 * np.abs(rec[:, 2].astype(np.float64)) * rec[:, 3].astype(np.float64))
    np.testing.assert_array_equal(got["wy_a"], wy.astype(np.float32))
    np.testing.assert_array_equal(got["wx_a"], wx.astype(np.float32))
    np.testing.assert_array_equal(got["wy_c"], wy)
    for key in ("a", "b"):
        exp = ref_width_f32(got[f"wx_{key}"], ny, got[f"wy_{key}"])
        assert np.array_equal(got[f"nx_{key}"], exp), key
    exp64 = ref_width_f64(got["wx_c"], ny, got["wy_c"])
    assert np.array_equal(got["nx_c"], exp64)
