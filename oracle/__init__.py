"""CPU oracle of the SAR focusing path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this package, and only as the reference results that GPU outputs are compared
with (or as the timed CPU baseline).  The
product path (uncc_sar_focusing_amd + libsarf.so) never imports it.

Numerics live in sarf_oracle.c (C, FP64 and an FP32-faithful restatement of
src/cpuBackProjection/cpuBackProjection.hpp); this module loads the compiled
library and adds numpy wrappers plus the numpy restatement of
matlab/PGA.m:137-239 (pga_f64).

Pinning (tests/test_oracle.py): FFT vs numpy.fft and a direct DFT; range
compression vs the closed-form Dirichlet kernel of a point target; BP vs the
matched-filter brute force, the SPIE10 point-target scenario peaks
(matlab/formImagePointTargetBP.m), the known peak value amp*Np*K/Nfft and
pulse-split additivity; entropy vs closed forms; the adjoint vs central
finite differences of the FP64 forward model; PGA vs an injected
polynomial phase error.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_DIR, "build", "libsarf_oracle.so")
CLIGHT = 299792458.0

_P = C.c_void_p
_i = C.c_int
_d = C.c_double
_f = C.c_float

_SIG = {
    "orc_num_threads": (_i, []),
    "orc_set_num_threads": (None, [_i]),
    "orc_fft_f64": (_i, [_P, _i, _i]),
    "orc_fft_f32": (_i, [_P, _i, _i]),
    "orc_range_bins_f64": (None, [_i, _d, _P]),
    "orc_range_compress_f64": (_i, [_i, _i, _i, _P, _P, _P, _P, _P]),
    "orc_range_compress_f32": (_i, [_i, _i, _i, _P, _P]),
    "orc_bp_f64": (_i, [_i, _i, _i, _i, _i, _i, _d, _d, _d, _d, _d, _P, _P, _P, _P, _i, _P]),
    "orc_bp_f32_faithful": (_i, [_i, _i, _i, _i, _i, _i, _f, _f, _f, _f, _f,
                                 _P, _P, _P, _P, _P, _P, _P]),
    "orc_mf_f64": (_i, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _P, _P, _P, _P, _P]),
    "orc_synth_ph_f64": (_i, [_i, _i, _i, _P, _P, _P, _P, _P, _d, _P]),
    "orc_entropy_f64": (_d, [_P, C.c_long]),
    "orc_bp_adjoint_f64": (_i, [_i, _i, _i, _i, _d, _d, _d, _d, _d, _P, _P, _P, _P, _P, _P]),
    "orc_db8": (_i, [_i, _i, _P, _f, _P]),
    "orc_db8_level_mismatches": (C.c_long, [_f, _P, C.POINTER(C.c_uint)]),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", _DIR, "--no-print-directory"], check=True,
                           stdout=subprocess.DEVNULL)
        _LIB = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(_LIB, name)
            fn.restype = res
            fn.argtypes = args
    return _LIB


def _p(a: np.ndarray) -> _P:
    return a.ctypes.data_as(_P)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int) -> None:
    lib().orc_set_num_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


def fft_f64(x, sign: int) -> np.ndarray:
    y = _c128(x).copy()
    if lib().orc_fft_f64(_p(y), y.size, sign) != 0:
        raise ValueError("length must be a power of two")
    return y


def range_bins(nfft: int, maxWr: float) -> np.ndarray:
    r = np.empty(nfft)
    lib().orc_range_bins_f64(nfft, maxWr, _p(r))
    return r


def range_compress_f64(ph, nfft: int, taper=None, ramp=None) -> np.ndarray:
    """ph (Np, K) -> prof (Np, nfft) = fftshift(ifft(ph, nfft)) [* ramp]."""
    ph = _c128(ph)
    Np, K = ph.shape
    out = np.empty((Np, nfft), dtype=np.complex128)
    tp = _f64(taper) if taper is not None else None
    f0 = r0 = None
    if ramp is not None:
        f0, r0 = _f64(ramp[0]), _f64(ramp[1])
    rc = lib().orc_range_compress_f64(K, Np, nfft, _p(ph), _p(tp) if tp is not None else None,
                                      _p(f0) if f0 is not None else None,
                                      _p(r0) if r0 is not None else None, _p(out))
    if rc != 0:
        raise ValueError("bad range-compression sizes")
    return out


def range_compress_f32(ph, nfft: int) -> np.ndarray:
    ph = np.ascontiguousarray(ph, dtype=np.complex64)
    Np, K = ph.shape
    out = np.empty((Np, nfft), dtype=np.complex64)
    if lib().orc_range_compress_f32(K, Np, nfft, _p(ph), _p(out)) != 0:
        raise ValueError("bad range-compression sizes")
    return out


def bp_f64(prof, ant, r0, f0, grid, maxWr: float, pulses=None, img=None) -> np.ndarray:
    """FP64 backprojection; returns img (Nx, Ny) complex128 (img[x, y])."""
    prof = _c128(prof)
    Np, nfft = prof.shape
    ant, r0, f0 = _f64(ant), _f64(r0), _f64(f0)
    acc = 1 if img is not None else 0
    if img is None:
        img = np.empty((grid.Nx, grid.Ny), dtype=np.complex128)
    b, e = (0, Np) if pulses is None else pulses
    rc = lib().orc_bp_f64(grid.Nx, grid.Ny, nfft, Np, b, e, grid.x0, grid.y0, grid.Wx, grid.Wy,
                          maxWr, _p(prof), _p(ant), _p(r0), _p(f0), acc, _p(img))
    if rc != 0:
        raise ValueError("bad backprojection arguments")
    return img


def bp_f32_faithful(prof, ant, r0, f0, grid, maxWr: float, pulses=None) -> np.ndarray:
    """FP32 restatement with the reference's float-isms (F1/F2)."""
    prof = np.ascontiguousarray(prof, dtype=np.complex64)
    Np, nfft = prof.shape
    a32 = np.asarray(ant, dtype=np.float32)
    ax, ay, az = (np.ascontiguousarray(a32[:, i]) for i in range(3))
    r0 = np.ascontiguousarray(r0, dtype=np.float32)
    f0 = np.ascontiguousarray(f0, dtype=np.float32)
    img = np.empty((grid.Nx, grid.Ny), dtype=np.complex64)
    b, e = (0, Np) if pulses is None else pulses
    rc = lib().orc_bp_f32_faithful(grid.Nx, grid.Ny, nfft, Np, b, e, grid.x0, grid.y0, grid.Wx,
                                   grid.Wy, maxWr, _p(prof), _p(ax), _p(ay), _p(az), _p(r0),
                                   _p(f0), _p(img))
    if rc != 0:
        raise ValueError("bad backprojection arguments")
    return img


def mf_f64(ph, ant, r0, f0, df: float, grid) -> np.ndarray:
    ph = _c128(ph)
    Np, K = ph.shape
    ant, r0, f0 = _f64(ant), _f64(r0), _f64(f0)
    img = np.empty((grid.Nx, grid.Ny), dtype=np.complex128)
    lib().orc_mf_f64(grid.Nx, grid.Ny, K, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, df, _p(ph),
                     _p(ant), _p(r0), _p(f0), _p(img))
    return img


def synth_ph_f64(geom, targets) -> np.ndarray:
    """Point-target phase history (Np, K) complex128."""
    pos = _f64(targets.pos)
    amp = _c128(targets.amp)
    ant, r0, f0 = _f64(geom.ant), _f64(geom.r0), _f64(geom.f0)
    Np, K = geom.Np, geom.K
    ph = np.empty((Np, K), dtype=np.complex128)
    lib().orc_synth_ph_f64(K, Np, pos.shape[0], _p(pos), _p(amp), _p(ant), _p(r0), _p(f0),
                           geom.df, _p(ph))
    return ph


def db8(img, dyn_range_db: float = 70.0) -> np.ndarray:
    """8-bit dB display image (Ny, Nx), top row = far range, as the
    reference's writeBMPFile quantises (uncc_sar_globals.hpp:66-118)."""
    img = np.ascontiguousarray(img, dtype=np.complex64)
    Nx, Ny = img.shape
    out = np.empty((Ny, Nx), dtype=np.uint8)
    if lib().orc_db8(Nx, Ny, _p(img), dyn_range_db, _p(out)) != 0:
        raise ValueError("bad db8 arguments")
    return out


def db8_level_mismatches(dyn_range_db: float, levels) -> tuple[int, int]:
    """Exhaustive check of a 256-entry level table against the reference's
    dB quantisation of r = |I|/max over every float r in [0, 1]: returns
    (number of mismatching r, smallest mismatching bit pattern or
    0xffffffff)."""
    lv = np.ascontiguousarray(levels, dtype=np.float32)
    if lv.shape != (256,):
        raise ValueError("levels must hold 256 floats")
    first = C.c_uint(0)
    bad = lib().orc_db8_level_mismatches(dyn_range_db, _p(lv), C.byref(first))
    return int(bad), int(first.value)


def reference_trajectory_f64(params, D: int, Np: int) -> np.ndarray:
    """(Np, 3) candidate phase centres of the reference's grid-search model
    (src/autofocus/gridSearchErrorFunctions_dft.cuh:279-291), pulse index
    i = 0..Np-1, evaluated in FP64:
      D = 10: x_i = p2 i^2 + p1 i p9 + p0, y_i = p5 i^2 + p4 i p9 + p3,
              z_i = p8 i^2 + p7 i p9 + p6;
      D = 7:  x_i = p1 i p6 + p0, y_i = p3 i p6 + p2, z_i = p5 i p6 + p4.
    Pinned by construction against those lines (tests/test_workloads.py
    checks the engine's linear form against it)."""
    p = np.asarray(params, dtype=np.float64)
    i = np.arange(Np, dtype=np.float64)
    if D == 10:
        return np.stack([p[3 * a + 2] * i * i + p[3 * a + 1] * i * p[9] + p[3 * a]
                         for a in range(3)], axis=1)
    if D == 7:
        return np.stack([p[2 * a + 1] * i * p[6] + p[2 * a] for a in range(3)], axis=1)
    raise ValueError("D must be 7 or 10")


def entropy_f64(img) -> float:
    img = _c128(img)
    return lib().orc_entropy_f64(_p(img), img.size)


def bp_adjoint_f64(prof, ant, r0, f0, grid, maxWr: float, img) -> np.ndarray:
    """dE/d(antenna positions), (Np, 3)."""
    prof = _c128(prof)
    Np, nfft = prof.shape
    ant, r0, f0 = _f64(ant), _f64(r0), _f64(f0)
    img = _c128(img)
    g = np.empty((Np, 3))
    rc = lib().orc_bp_adjoint_f64(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx,
                                  grid.Wy, maxWr, _p(prof), _p(ant), _p(r0), _p(f0), _p(img),
                                  _p(g))
    if rc != 0:
        raise ValueError("adjoint: empty image")
    return g


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---------------------------------------------------------------------------
# PGA, restated from matlab/PGA.m:137-239 (numpy; FFTs are numpy.fft).
# Image convention: MATLAB img is Ny x Nx (rows = range y, columns = azimuth
# x).  Our arrays are img[x, y], so MATLAB row r is our img[:, r] and the
# "along dimension 2" operations act on axis 0 here.
# ---------------------------------------------------------------------------
def pga_f64(img_xy, iters: int, db_down: float = 30.0):
    """Returns (corrected image img[x, y], accumulated phase af_ph (Nx,),
    per-iteration rms of the detrended estimate)."""
    A = np.array(img_xy, dtype=np.complex128).T.copy()  # MATLAB layout (Ny, Nx)
    Ny, Nx = A.shape
    center = int(np.ceil(Nx / 2))  # 1-based centre index (PGA.m:151)
    af_ph = np.zeros(Nx)
    rms = []
    for _ in range(iters):
        # brightest azimuth sample per range row, circularly shifted so it
        # lands at 1-based column center + 1 (PGA.m:159-164)
        idx = np.argmax(np.abs(A), axis=1) + 1  # 1-based
        f = np.empty_like(A)
        for r in range(Ny):
            f[r] = np.roll(A[r], 1 + center - idx[r])
        # noncoherent window from the unshifted image (PGA.m:167-188)
        w = np.sum(np.conj(A) * A, axis=0).real
        wdb = 10 * np.log10(w / w.max())
        cut = wdb.max() - abs(db_down)
        left_c = np.nonzero(wdb[:center] - cut < 0)[0]
        right_c = np.nonzero(wdb[center:] - cut < 0)[0]
        left = (left_c[-1] + 1) + 1 if left_c.size else 1          # 1-based
        right = (right_c[0] + 1) + center - 1 if right_c.size else Nx
        mask = np.zeros(Nx)
        mask[left - 1:right] = 1.0
        g = f * mask[None, :]
        # phase-gradient estimate (PGA.m:195-204)
        G = np.fft.fft(np.fft.ifftshift(g, axes=1), axis=1)
        Gd = np.diff(G, axis=1)
        num = np.sum(np.imag(np.conj(G[:, :-1]) * Gd), axis=0)
        den = np.sum(np.abs(G[:, :-1]) ** 2, axis=0)
        phi_dot = num / den
        phi = np.unwrap(np.concatenate([[0.0], np.cumsum(phi_dot)]))
        # remove the linear trend (PGA.m:209-212)
        x = np.arange(Nx, dtype=np.float64)
        c1, c0 = np.polyfit(x, phi, 1)
        phi_d = phi - (c1 * x + c0)
        rms.append(float(np.sqrt(np.mean(phi_d ** 2))))
        af_ph += phi_d
        # apply (PGA.m:229-232)
        IMG = np.fft.fft(A, axis=1) * np.exp(-1j * phi_d)[None, :]
        A = np.fft.ifft(IMG, axis=1)
    return A.T.copy(), af_ph, rms
