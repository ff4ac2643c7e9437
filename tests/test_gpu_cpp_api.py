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
CLIGHT = 299792458.0


def _sandia_record(tmp_path, K=424, Np=128, targets=((1.0, -2.0, 1.0), (-3.0, 4.0, 0.7),
                                                    (6.0, 1.5, 0.5))):
    """A Sandia-format MAT record (scipy-written, compressed, 2 channels: HH
    and HV = 0.5 HH) of point targets (x, y, amplitude).  Returns the path
    and the float32 values readData will hand to the callback (channel HV)."""
    import scipy.io

    adf = 1.0e6
    f0, bw = 9.6e9, 640e6
    df = bw / (K - 1)
    th = np.deg2rad(90.0 - 2.0 + 4.0 * np.arange(Np) / (Np - 1))
    pos = np.stack([10000.0 * np.cos(th), 10000.0 * np.sin(th), np.full(Np, 7500.0)])
    pos = pos.astype(np.float32).astype(np.float64)  # what the float reader will see
    r0 = np.sqrt((pos ** 2).sum(0))
    fk = f0 + df * np.arange(K)
    ph = np.zeros((K, Np), dtype=np.complex128)
    for tx, ty, a in targets:
        dR = np.sqrt((pos[0] - tx) ** 2 + (pos[1] - ty) ** 2 + pos[2] ** 2) - r0
        ph += a * np.exp(-4j * np.pi * fk[:, None] * dR[None, :] / CLIGHT)  # K x Np
    ph = ph.astype(np.complex64)
    sd = np.stack([ph, (0.5 * ph).astype(np.complex64)], axis=2)
    start = np.full((1, Np), f0, dtype=np.float32)
    chirp = np.full((1, Np), df * adf, dtype=np.float32)
    rec = {"sph_MATData": {"preamble": {"ADF": np.array([[adf]], dtype=np.float32)},
                           "Data": {"SampleData": sd, "StartF": start, "ChirpRate": chirp,
                                    "radarCoordinateFrame": {"x": pos[0:1], "y": pos[1:2],
                                                             "z": pos[2:3]}}}}
    mat = tmp_path / "scene.mat"
    scipy.io.savemat(mat, rec, do_compression=True)
    return mat, dict(ph=(0.5 * ph).astype(np.complex64).T.copy(), pos=pos.T.copy(),
                     startF=start[0], chirp=chirp[0], adf=np.float32(adf), K=K, Np=Np,
                     targets=targets)


def _ref_width_f32(wx, ny, wy):
    """The reference's default width (int)((float)Wx * N_y) / Wy in float
    (include/uncc_sar_focusing.hpp:461, charlotte_sar_api.hpp:189)."""
    num = int(np.trunc(np.float32(wx) * np.float32(ny)))
    return int(np.trunc(np.float32(num) / np.float32(wy)))


def test_read_focus_sandia_mat_end_to_end(tmp_path):
    """readData -> sph_sar_data_callback_gpu -> BMP on a Sandia record with
    K = 424 samples (not a power of two: the callback's N_fft = K).  The
    default image parameters equal the reference's formulas (N_fft = N_y = K,
    max_Wy = c / (2 mean deltaF), N_x = (int)((float)Wx N_y) / Wy), the image
    the callback forms equals the FP64 oracle's range compression +
    backprojection of the same record (rel-L2 <= 1e-4, same brightest pixel,
    brightest pixel on the strongest target), and the BMP the callback wrote
    is byte-identical to the reference writer's quantisation of that image."""
    if not os.path.exists(RF):
        pytest.skip("read_focus not built")
    mat, r = _sandia_record(tmp_path)
    K, Np = r["K"], r["Np"]
    bmp, raw = tmp_path / "scene.bmp", tmp_path / "scene.bin"
    p = subprocess.run([RF, "--image", str(mat), str(bmp), str(raw), "100000", "HV"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    buf = raw.read_bytes()
    Nx, Ny, nfft, k2, np2 = np.frombuffer(buf[:20], dtype="<i4")
    x0, y0, Wx, Wy, maxWy = np.frombuffer(buf[20:60], dtype="<f8")
    r0 = np.frombuffer(buf[60:60 + 8 * Np], dtype="<f8")
    f0 = np.frombuffer(buf[60 + 8 * Np:60 + 16 * Np], dtype="<f8")
    img = np.frombuffer(buf[60 + 16 * Np:], dtype=np.complex64).reshape(Nx, Ny)
    assert (k2, np2) == (K, Np)
    # image parameters of create_SARImageFormationParams (float aperture)
    df32 = np.float32(r["chirp"][0] / r["adf"])
    mean_df = np.float32(np.float64(df32))  # mean of identical per-pulse steps
    assert nfft == K and Ny == K
    assert np.float32(maxWy) == np.float32(CLIGHT / (2.0 * np.float64(mean_df)))
    assert Wy == maxWy and x0 == 0.0 and y0 == 0.0
    assert Nx == _ref_width_f32(Wx, Ny, Wy)
    # aperture fields: R0 = |a| rounded to float, startF as recorded
    np.testing.assert_array_equal(r0, np.sqrt((r["pos"] ** 2).sum(1)).astype(np.float32))
    np.testing.assert_array_equal(f0, r["startF"].astype(np.float64))
    # the image against the FP64 oracle on the same record
    from uncc_sar_focusing_amd.scenes import ImageGrid

    grid = ImageGrid(int(Nx), int(Ny), float(x0), float(y0), float(Wx), float(Wy))
    prof = oracle.range_compress_f64(r["ph"], int(nfft))
    ref = oracle.bp_f64(prof, r["pos"], r0, f0, grid, float(maxWy))
    err = oracle.rel_l2(img, ref)
    print(f"default image {Nx}x{Ny}, N_fft {nfft}: rel-L2 vs FP64 oracle {err:.3e}")
    assert err <= 1e-4
    assert np.argmax(np.abs(img)) == np.argmax(np.abs(ref))
    bx, by = np.unravel_index(int(np.argmax(np.abs(img))), img.shape)
    tx, ty, _ = r["targets"][0]
    assert abs(grid.x()[bx] - tx) <= 1.5 * Wx / (Nx - 1)
    assert abs(grid.y()[by] - ty) <= 1.5 * Wy / (Ny - 1)
    # the callback's BMP is the reference writer's file of this image
    assert bmp.read_bytes() == encode_bmp(oracle.db8(img, 70.0))


def test_focus_sar_image_matched_filter_matches_oracle(tmp_path):
    """focus_SAR_image with algorithm MATCHED_FILTER (the reference's CPU
    entry dispatches it to run_mf, cpuBackProjection.hpp:255-257) on the
    same Sandia record: equals the FP64 oracle matched filter over pulses
    2..Np-1 (run_mf starts at pulse 2, :218) within 1e-4 rel-L2, with the
    same brightest pixel; including pulses 0 and 1 would differ."""
    if not os.path.exists(RF):
        pytest.skip("read_focus not built")
    mat, r = _sandia_record(tmp_path, K=200, Np=48)
    K = r["K"]
    raw = tmp_path / "mf.bin"
    Nx, Ny, W = 48, 40, 12.0
    p = subprocess.run([RF, "--mf", str(mat), str(raw), str(Nx), str(Ny), str(W), str(W),
                        "100000", "HV"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    buf = raw.read_bytes()
    assert tuple(np.frombuffer(buf[:8], dtype="<i4")) == (Nx, Ny)
    img = np.frombuffer(buf[8:], dtype=np.complex64).reshape(Nx, Ny)
    from uncc_sar_focusing_amd.scenes import ImageGrid

    grid = ImageGrid(Nx, Ny, 0.0, 0.0, np.float32(W), np.float32(W))
    # aperture frequencies: freq[n,k] = startF + k * (chirpRate / ADF) in float
    df32 = np.float32(r["chirp"][0] / r["adf"])
    f_first = np.float32(r["startF"][0])
    f_last = np.float32(f_first + np.float32(np.float32(K - 1) * df32))
    df = (np.float64(f_last) - np.float64(f_first)) / (K - 1)
    r0 = np.sqrt((r["pos"] ** 2).sum(1)).astype(np.float32).astype(np.float64)
    f0 = np.full(r["Np"], np.float64(f_first))
    ref = oracle.mf_f64(r["ph"][2:], r["pos"][2:], r0[2:], f0[2:], df, grid)
    err = oracle.rel_l2(img, ref)
    print(f"matched filter {Nx}x{Ny}: rel-L2 vs FP64 oracle (pulses 2..) {err:.3e}")
    assert err <= 1e-4
    assert np.argmax(np.abs(img)) == np.argmax(np.abs(ref))
    every = oracle.mf_f64(r["ph"], r["pos"], r0, f0, df, grid)
    assert oracle.rel_l2(img, every) > 1e-3
