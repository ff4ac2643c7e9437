"""Host-side check of the parallel copier used by the host-buffer focus path
(csrc/sarf_hostcopy.hpp): results equal memcpy for random sizes and offsets,
with and without ThreadSanitizer.  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "uncc_sar_focusing_amd", "csrc")


def _build_and_run(tmp_path, extra):
    exe = str(tmp_path / "hostcopy_check")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", *extra, f"-I{CSRC}",
                    os.path.join(HERE, "host", "hostcopy_check.cpp"), "-o", exe], check=True)
    return subprocess.run([exe], capture_output=True, text=True, timeout=300)


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_hostcopy_matches_memcpy(tmp_path):
    r = _build_and_run(tmp_path, [])
    assert r.returncode == 0, r.stderr
    assert "hostcopy ok" in r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_hostcopy_under_thread_sanitizer(tmp_path):
    r = _build_and_run(tmp_path, ["-fsanitize=thread"])
    assert r.returncode == 0, r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr
