"""The level table behind the GPU dB quantisation (sarf_db8_levels, host code
of libsarf) reproduces the reference's float expression
(include/uncc_sar_globals.hpp:99-104: (float)(255/dyn) * (20 log10(r) + dyn),
truncated to a byte, negative -> 0) for EVERY float r = |I|/max in [0, 1]:
the oracle evaluates the expression with the host's log10f on all
1065353217 bit patterns and counts the levels at or below each.  A table
moved by one float at one level fails the same check.  CPU only (libsarf is
loaded, no device is touched)."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from uncc_sar_focusing_amd import _lib

pytestmark = pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="libsarf.so not built")


def levels(dyn):
    lv = np.zeros(256, np.float32)
    _lib.check(_lib.lib().sarf_db8_levels(dyn, lv.ctypes.data_as(ctypes.c_void_p)))
    return lv


@pytest.mark.parametrize("dyn", [70.0, 40.0, 100.5])
def test_levels_reproduce_reference_bytes_for_every_float(dyn):
    lv = levels(dyn)
    assert lv[0] == 0.0
    assert np.all(np.diff(lv[1:]) >= 0)
    bad, first = oracle.db8_level_mismatches(dyn, lv)
    assert bad == 0, f"{bad} mismatches, first at bits {first:#x}"


def test_check_detects_a_level_off_by_one_float():
    lv = levels(70.0)
    for b, d in ((100, -1), (200, +1), (1, +1)):
        moved = lv.copy()
        moved[b] = np.nextafter(moved[b], np.float32(d * 2.0))
        bad, first = oracle.db8_level_mismatches(70.0, moved)
        assert bad == 1, (b, d, bad)


def test_levels_reject_bad_range():
    out = np.zeros(256, np.float32)
    p = out.ctypes.data_as(ctypes.c_void_p)
    assert _lib.lib().sarf_db8_levels(0.0, p) == _lib.SARF_ERR_INVALID
    assert _lib.lib().sarf_db8_levels(float("nan"), p) == _lib.SARF_ERR_INVALID
    assert _lib.lib().sarf_db8_levels(70.0, None) == _lib.SARF_ERR_INVALID
