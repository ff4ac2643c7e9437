"""The C ABI library loads and exports every symbol include/sarf.h declares.

No compute calls: these run on CPU-only machines.
"""
import ctypes
import os
import re

import pytest

from uncc_sar_focusing_amd import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in sorted(os.listdir(os.path.join(REPO, "include")))
           if h.endswith(".h")]


def declared_symbols():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b(sarf_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_documented_entry_points():
    syms = declared_symbols()
    for s in ["sarf_ctx_create", "sarf_ctx_destroy", "sarf_last_error", "sarf_range_compress",
              "sarf_backproject", "sarf_focus", "sarf_version"]:
        assert s in syms


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libsarf.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() <= set(_lib.SIGNATURES), declared_symbols() - set(_lib.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libsarf.so not built")
def test_version_and_error_string_without_device():
    lib = _lib.load()
    assert lib.sarf_version() == 101
    assert lib.sarf_last_error() is not None


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libsarf.so not built")
def test_invalid_context_argument_is_an_error_code():
    """sarf_ctx_create(out=NULL) fails with SARF_ERR_INVALID before touching
    any device."""
    lib = _lib.load()
    rc = lib.sarf_ctx_create(0, None)
    assert rc == _lib.SARF_ERR_INVALID
    assert b"NULL" in lib.sarf_last_error()


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _lib.load(str(tmp_path / "nope.so"))


def test_cpp_interface_headers_compile_with_plain_gcc(tmp_path):
    """include/*.hpp (the reference's C++ surface over the C ABI) needs only a
    C++17 compiler and sarf.h -- no HIP headers -- to compile caller code."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = os.path.join(REPO, "uncc_sar_focusing_amd", "examples", "example_gpu.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(REPO, "include"), src],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
