"""MAT-file ingest (readData, charlotte_sar_api.cpp:36-148) through libsarf's
own Level 5 reader, on records written by scipy.io.savemat in the layouts
the reference reads (uncc_sar_matio.cpp:23-52): Sandia sph_MATData with 1 or
4 polarity channels, compressed or not, single or double precision, and
GOTCHA data.*.  CPU only: the --dump mode reads and writes, no GPU call."""
import os
import subprocess

import numpy as np
import pytest
import scipy.io

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "uncc_sar_focusing_amd", "examples", "read_focus")


def dump(path, tmp_path, max_pulses=1 << 30, pol="any"):
    out = tmp_path / "dump.bin"
    r = subprocess.run([EXE, "--dump", str(path), str(out), str(max_pulses), pol],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=60)
    if r.returncode != 0:
        return r.returncode, r.stdout, None
    b = out.read_bytes()
    K, Np = np.frombuffer(b[:8], dtype=np.int32)
    f = np.frombuffer(b[8:], dtype=np.float32)
    adf = f[0]
    sd = f[1:1 + 2 * K * Np].reshape(Np, K, 2)
    rest = f[1 + 2 * K * Np:].reshape(5, Np)
    return 0, r.stdout, dict(K=int(K), Np=int(Np), adf=float(adf), sd=sd[..., 0] + 1j * sd[..., 1],
                             startF=rest[0], chirp=rest[1], x=rest[2], y=rest[3], z=rest[4])


def sandia(K, Np, npol, dtype=np.float64, seed=1):
    rng = np.random.default_rng(seed)
    shape = (K, Np, npol) if npol > 1 else (K, Np)
    sd = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(
        np.complex64 if dtype == np.float32 else np.complex128)
    pp = lambda: rng.uniform(-1e4, 1e4, (1, Np)).astype(dtype)  # noqa: E731
    rec = {"sph_MATData": {
        "preamble": {"ADF": np.array([[2.5e6]], dtype=dtype)},
        "total_pulses": np.array([[Np]], dtype=np.float64),
        "Data": {"SampleData": sd,
                 "StartF": np.linspace(9.5e9, 9.6e9, Np).reshape(1, Np).astype(dtype),
                 "ChirpRate": np.full((1, Np), 4.0e12, dtype=dtype),
                 "radarCoordinateFrame": {"x": pp(), "y": pp(), "z": pp()},
                 "label": "a char field that is skipped"}}}
    return rec, sd


@pytest.fixture(autouse=True)
def _need_exe():
    if not os.path.exists(EXE):
        pytest.skip("read_focus not built")


@pytest.mark.parametrize("compress", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sandia_single_channel(tmp_path, compress, dtype):
    rec, sd = sandia(40, 30, 1, dtype)
    path = tmp_path / "sph.mat"
    scipy.io.savemat(path, rec, do_compression=compress)
    rc, log, d = dump(path, tmp_path)
    assert rc == 0, log
    assert (d["K"], d["Np"]) == (40, 30)
    assert d["adf"] == np.float32(2.5e6)
    np.testing.assert_array_equal(d["sd"], sd.T.astype(np.complex64))
    data = rec["sph_MATData"]["Data"]
    np.testing.assert_array_equal(d["startF"], data["StartF"][0].astype(np.float32))
    np.testing.assert_array_equal(d["chirp"], data["ChirpRate"][0].astype(np.float32))
    for ax in "xyz":
        np.testing.assert_array_equal(d[ax], data["radarCoordinateFrame"][ax][0].astype(np.float32))


def test_sandia_polarity_and_central_pulses(tmp_path):
    rec, sd = sandia(16, 11, 4)
    path = tmp_path / "sph4.mat"
    scipy.io.savemat(path, rec, do_compression=True)
    for ch, pol in enumerate(["HH", "HV", "VH", "VV"]):
        rc, log, d = dump(path, tmp_path, 6, pol)
        assert rc == 0, log
        # central 6 of 11 pulses: first = (11 - 6) // 2 = 2
        assert d["Np"] == 6
        np.testing.assert_array_equal(d["sd"], sd[:, 2:8, ch].T.astype(np.complex64))
        np.testing.assert_array_equal(
            d["x"], rec["sph_MATData"]["Data"]["radarCoordinateFrame"]["x"][0, 2:8].astype(np.float32))
    rec1, _ = sandia(16, 11, 1)
    p1 = tmp_path / "sph1.mat"
    scipy.io.savemat(p1, rec1)
    rc, log, _ = dump(p1, tmp_path, pol="VV")
    assert rc != 0 and "not available" in log


def test_gotcha_record(tmp_path):
    rng = np.random.default_rng(4)
    K, Np = 24, 9
    fp = rng.standard_normal((K, Np)) + 1j * rng.standard_normal((K, Np))
    freq = (9.6e9 + 1.5e6 * np.arange(K)).reshape(K, 1)
    rec = {"data": {"fp": fp, "freq": freq, "x": rng.uniform(-1, 1, (1, Np)) * 1e4,
                    "y": rng.uniform(-1, 1, (1, Np)) * 1e4, "z": np.full((1, Np), 7.5e3),
                    "r0": np.full((1, Np), 1.25e4), "th": np.zeros((1, Np)), "phi": np.zeros((1, Np)),
                    "af": {"r_correct": np.zeros((1, Np)), "ph_correct": np.zeros((1, Np))}}}
    path = tmp_path / "gotcha.mat"
    scipy.io.savemat(path, rec, do_compression=True)
    rc, log, d = dump(path, tmp_path)
    assert rc == 0, log
    np.testing.assert_array_equal(d["sd"], fp.T.astype(np.complex64))
    assert d["adf"] == 1.0
    np.testing.assert_allclose(d["startF"], np.float32(9.6e9))
    np.testing.assert_allclose(d["chirp"], 1.5e6, rtol=1e-6)
    np.testing.assert_array_equal(d["z"], np.float32(7.5e3))


def test_bad_files_fail_cleanly(tmp_path):
    p = tmp_path / "junk.mat"
    p.write_bytes(b"not a mat file" * 20)
    rc, log, _ = dump(p, tmp_path)
    assert rc != 0 and "readData" in log
    rc, log, _ = dump(tmp_path / "missing.mat", tmp_path)
    assert rc != 0 and "cannot open" in log
    good, _ = sandia(8, 5, 1)
    q = tmp_path / "trunc.mat"
    scipy.io.savemat(q, good, do_compression=True)
    q.write_bytes(q.read_bytes()[:-40])
    rc, log, _ = dump(q, tmp_path)
    assert rc != 0
