"""Python host interface of the MI355X SAR focusing engine.

Thin wrappers over the C ABI (include/sarf.h) that take and return torch
tensors resident on the GPU; torch is used only for device memory and the
current HIP stream.  Array conventions follow the reference
(src/cpuBackProjection/cpuBackProjection.hpp:87, charlotte_sar_api.hpp:145-148):

  phase history  (Np, K)     complex64, pulse-major, K samples per pulse
  range profiles (Np, Nfft)  complex64
  antenna        (Np, 3)     float64
  image          (Nx, Ny)    complex64 -- memory layout img[x*Ny + y]
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import BpDesc, FocusDesc, RcDesc, SarfError, check  # noqa: F401
from .scenes import ImageGrid

_P = C.c_void_p


def _ptr(t: torch.Tensor | None) -> _P:
    return _P(0 if t is None else t.data_ptr())


def _need(t: torch.Tensor, dtype: torch.dtype, shape: tuple, name: str) -> None:
    if not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class Context:
    """Owns a libsarf context bound to one GPU."""

    def __init__(self, device: int = 0):
        self.lib = _lib.lib()
        self.device = device
        h = C.c_void_p()
        check(self.lib.sarf_ctx_create(device, C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.sarf_ctx_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _stream(self) -> _P:
        return _P(torch.cuda.current_stream(self.device).cuda_stream)

    # ------------------------------------------------------------------ RC
    def range_compress(self, ph: torch.Tensor, nfft: int, *, taper: torch.Tensor | None = None,
                       ramp: tuple[torch.Tensor, torch.Tensor] | None = None,
                       out: torch.Tensor | None = None) -> torch.Tensor:
        """fftshift(ifft(ph, nfft)) per pulse (bpBasic.m:82).  complex64 or
        complex128 (FP64 path)."""
        Np, K = ph.shape
        f64 = ph.dtype == torch.complex128
        cdt = torch.complex128 if f64 else torch.complex64
        _need(ph, cdt, (Np, K), "ph")
        flags = _lib.SARF_RC_F64 if f64 else 0
        if taper is not None:
            _need(taper, torch.float32, (K,), "taper")
            flags |= _lib.SARF_RC_TAPER
        f0 = r0 = None
        if ramp is not None:
            f0, r0 = ramp
            _need(f0, torch.float64, (Np,), "f0")
            _need(r0, torch.float64, (Np,), "r0")
            flags |= _lib.SARF_RC_RAMP
        if out is None:
            out = torch.empty((Np, nfft), dtype=cdt, device=ph.device)
        _need(out, cdt, (Np, nfft), "out")
        d = RcDesc(K, Np, nfft, flags)
        check(self.lib.sarf_range_compress(self._h, C.byref(d), _ptr(ph), _ptr(taper), _ptr(f0),
                                           _ptr(r0), _ptr(out), self._stream()))
        return out

    def range_compress_scenes(self, phs, nffts, *, tapers=None, ramps=None,
                              out=None) -> list:
        """range_compress of many complex64 phase histories (shapes and nfft
        may differ) in one call, one launch per transform length
        (sarf_range_compress_scenes).  `tapers` / `ramps` optionally give a
        per-scene taper tensor / (f0, r0) pair or None; `out` may pass the
        profile tensors to write (for graph capture)."""
        n = len(phs)
        if len(nffts) != n:
            raise ValueError("range_compress_scenes: one nfft per phase history")
        descs = (_lib.RcScene * n)()
        res = []
        for i, ph in enumerate(phs):
            Np, K = ph.shape
            nfft = int(nffts[i])
            _need(ph, torch.complex64, (Np, K), f"scene {i} ph")
            flags = 0
            taper = tapers[i] if tapers is not None else None
            if taper is not None:
                _need(taper, torch.float32, (K,), f"scene {i} taper")
                flags |= _lib.SARF_RC_TAPER
            f0 = r0 = None
            if ramps is not None and ramps[i] is not None:
                f0, r0 = ramps[i]
                _need(f0, torch.float64, (Np,), f"scene {i} f0")
                _need(r0, torch.float64, (Np,), f"scene {i} r0")
                flags |= _lib.SARF_RC_RAMP
            prof = out[i] if out is not None else torch.empty(
                (Np, nfft), dtype=torch.complex64, device=ph.device)
            _need(prof, torch.complex64, (Np, nfft), f"scene {i} out")
            descs[i] = _lib.RcScene(*(0 if t is None else t.data_ptr()
                                      for t in (ph, prof, taper, f0, r0)), K, Np, nfft, flags)
            res.append(prof)
        check(self.lib.sarf_range_compress_scenes(self._h, n, descs, self._stream()))
        return res

    # ------------------------------------------------------------------ BP
    def backproject(self, prof: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor,
                    f0: torch.Tensor, grid: ImageGrid, maxWr: float, *,
                    precision: str = "mixed", out: torch.Tensor | None = None,
                    accumulate: bool = False, tile: int = 0,
                    pulse_split: int = 0, los: tuple[float, float] | None = None) -> torch.Tensor:
        """Backproject profiles onto `grid`.  `los` optionally gives the
        horizontal unit line of sight (x, y) from the image centre to the
        aperture centre; it only steers the default tile shape."""
        Np, nfft = prof.shape
        f64 = precision == "f64"
        cdt = torch.complex128 if f64 else torch.complex64
        _need(prof, cdt, (Np, nfft), "prof")
        _need(ant, torch.float64, (Np, 3), "ant")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        if out is None:
            alloc = torch.zeros if accumulate else torch.empty
            out = alloc((grid.Nx, grid.Ny), dtype=cdt, device=prof.device)
        _need(out, cdt, (grid.Nx, grid.Ny), "out")
        d = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                   _lib.SARF_BP_F64 if f64 else _lib.SARF_BP_MIXED,
                   _lib.SARF_BP_ACCUMULATE if accumulate else 0, tile, pulse_split,
                   *(los if los is not None else (0.0, 0.0)))
        check(self.lib.sarf_backproject(self._h, C.byref(d), _ptr(prof), _ptr(ant), _ptr(r0),
                                        _ptr(f0), _ptr(out), self._stream()))
        return out

    def matched_filter(self, ph: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor,
                       f0: torch.Tensor, df: torch.Tensor | float, grid: ImageGrid, *,
                       out: torch.Tensor | None = None, accumulate: bool = False,
                       pulse_split: int = 0) -> torch.Tensor:
        """Exact matched-filter image (run_mf, cpuBackProjection.hpp:165-228):
        sum over pulses and samples of ph * exp(j 4 pi (f0 + k df) dR / c).
        `df` is the per-pulse frequency step (tensor (Np,)) or one value."""
        Np, K = ph.shape
        _need(ph, torch.complex64, (Np, K), "ph")
        _need(ant, torch.float64, (Np, 3), "ant")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        if not isinstance(df, torch.Tensor):
            df = torch.full((Np,), float(df), dtype=torch.float64, device=ph.device)
        _need(df, torch.float64, (Np,), "df")
        if out is None:
            alloc = torch.zeros if accumulate else torch.empty
            out = alloc((grid.Nx, grid.Ny), dtype=torch.complex64, device=ph.device)
        _need(out, torch.complex64, (grid.Nx, grid.Ny), "out")
        d = _lib.MfDesc(grid.Nx, grid.Ny, K, Np, grid.x0, grid.y0, grid.Wx, grid.Wy,
                        _lib.SARF_BP_ACCUMULATE if accumulate else 0, pulse_split)
        check(self.lib.sarf_matched_filter(self._h, C.byref(d), _ptr(ph), _ptr(ant), _ptr(r0),
                                           _ptr(f0), _ptr(df), _ptr(out), self._stream()))
        return out

    def image_db8(self, img: torch.Tensor, dyn_range_db: float = 70.0,
                  out: torch.Tensor | None = None) -> torch.Tensor:
        """8-bit dB display image (Ny, Nx) uint8 on the device, top row = far
        range, quantised as the reference's writeBMPFile does
        (uncc_sar_globals.hpp:66-118)."""
        Nx, Ny = img.shape
        _need(img, torch.complex64, (Nx, Ny), "img")
        if out is None:
            out = torch.empty((Ny, Nx), dtype=torch.uint8, device=img.device)
        _need(out, torch.uint8, (Ny, Nx), "out")
        check(self.lib.sarf_image_db8(self._h, _ptr(img), Nx, Ny, float(dyn_range_db), _ptr(out),
                                      self._stream()))
        return out

    def autotune(self, prof: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor,
                 f0: torch.Tensor, grid: ImageGrid, maxWr: float, *, reps: int = 3,
                 out: torch.Tensor | None = None,
                 los: tuple[float, float] | None = None) -> tuple[dict, torch.Tensor]:
        """Time every accurate (tile, pulse split) pair for this geometry on
        these inputs; later `backproject` calls with tile=0, pulse_split=0
        and the same geometry use the fastest.  Returns ({tile, pulse_split,
        ms, default_ms, candidates}, image made with the chosen pair)."""
        Np, nfft = prof.shape
        _need(prof, torch.complex64, (Np, nfft), "prof")
        _need(ant, torch.float64, (Np, 3), "ant")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        if out is None:
            out = torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64, device=prof.device)
        _need(out, torch.complex64, (grid.Nx, grid.Ny), "out")
        d = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                   _lib.SARF_BP_MIXED, 0, 0, 0, *(los if los is not None else (0.0, 0.0)))
        r = _lib.TuneResult()
        check(self.lib.sarf_autotune(self._h, C.byref(d), _ptr(prof), _ptr(ant), _ptr(r0),
                                     _ptr(f0), _ptr(out), reps, C.byref(r), self._stream()))
        res = {k: getattr(r, k) for k, _ in _lib.TuneResult._fields_}
        return res, out

    def tune_table(self) -> list:
        """The launch shapes `autotune` remembered on this context, as a list
        of dicts (geometry: Nx, Ny, Np, Nfft, dx, dy, maxWr, los_x, los_y;
        choice: tile, pulse_split) that `json.dump` can write."""
        n = C.c_int(0)
        check(self.lib.sarf_tune_table_get(self._h, None, 0, C.byref(n)))
        arr = (_lib.TuneEntry * max(n.value, 1))()
        check(self.lib.sarf_tune_table_get(self._h, arr, n.value, C.byref(n)))
        return [{k: getattr(arr[i], k) for k, _ in _lib.TuneEntry._fields_}
                for i in range(n.value)]

    def load_tune_table(self, entries) -> None:
        """Adds the shapes of `entries` (as `tune_table` returns them, e.g.
        read back with `json.load`) to this context; later `backproject`
        calls with tile=0, pulse_split=0 on a matching geometry use them."""
        entries = list(entries)
        arr = (_lib.TuneEntry * max(len(entries), 1))()
        for i, e in enumerate(entries):
            arr[i] = _lib.TuneEntry(*(e[k] for k, _ in _lib.TuneEntry._fields_))
        check(self.lib.sarf_tune_table_set(self._h, arr, len(entries)))

    # ------------------------------------------------------------- entropy
    def image_entropy(self, img: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Device tensor [E, S, T] (float64): E = -sum p ln p, p = |I|^2/S."""
        if img.dtype not in (torch.complex64, torch.complex128) or not img.is_cuda:
            raise ValueError("img must be a complex CUDA tensor")
        if not img.is_contiguous():
            raise ValueError("img must be contiguous")
        if out is None:
            out = torch.empty(3, dtype=torch.float64, device=img.device)
        check(self.lib.sarf_image_entropy(self._h, _ptr(img), img.numel(),
                                          1 if img.dtype == torch.complex128 else 0, _ptr(out),
                                          self._stream()))
        return out

    def batch_entropy(self, prof: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor,
                      f0: torch.Tensor, grid: ImageGrid, maxWr: float, coef: torch.Tensor,
                      basis: torch.Tensor, direction: torch.Tensor, *, tile: int = 0,
                      los: tuple[float, float] | None = None,
                      out: torch.Tensor | None = None) -> torch.Tensor:
        """Entropy of the image of every candidate trajectory
        a_n + (coef[c] . basis[:, n]) * direction[n], one persistent launch.
        `direction` may instead be (nbasis, Np, 3), one field per basis:
        a_n + sum_k coef[c, k] basis[k, n] direction[k, n]."""
        Np, nfft = prof.shape
        ncand, nb = coef.shape
        _need(prof, torch.complex64, (Np, nfft), "prof")
        _need(ant, torch.float64, (Np, 3), "ant")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        _need(coef, torch.float64, (ncand, nb), "coef")
        _need(basis, torch.float64, (nb, Np), "basis")
        per_basis = direction.dim() == 3
        _need(direction, torch.float64, (nb, Np, 3) if per_basis else (Np, 3), "direction")
        if out is None:
            out = torch.empty(ncand, dtype=torch.float64, device=prof.device)
        _need(out, torch.float64, (ncand,), "out")
        d = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                   _lib.SARF_BP_MIXED, _lib.SARF_BATCH_DIR_PER_BASIS if per_basis else 0, tile, 0,
                   *(los if los is not None else (0.0, 0.0)))
        check(self.lib.sarf_bp_batch_entropy(self._h, C.byref(d), _ptr(prof), _ptr(ant),
                                             _ptr(r0), _ptr(f0), ncand, nb, _ptr(coef),
                                             _ptr(basis), _ptr(direction), _ptr(out),
                                             self._stream()))
        return out

    def grid_search(self, prof: torch.Tensor, r0: torch.Tensor, f0: torch.Tensor,
                    grid: ImageGrid, maxWr: float, start, end, num, levels: int, *,
                    model: str = "coef", ant: torch.Tensor | None = None,
                    basis: torch.Tensor | None = None, direction: torch.Tensor | None = None,
                    shrink: str = "reference", tile: int = 0,
                    los: tuple[float, float] | None = None):
        """Multi-resolution grid search (sarf_grid_search).  model "coef":
        parameters are batch_entropy coefficients (needs ant, basis,
        direction); "ref7" / "ref10": the reference's absolute trajectory
        parameters (ant/basis/direction unused).  Returns (best (dims,),
        best_entropy, covariance (dims, dims), per-level dicts)."""
        import numpy as np

        Np, nfft = prof.shape
        _need(prof, torch.complex64, (Np, nfft), "prof")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        start = np.asarray(start, dtype=np.float64)
        end = np.asarray(end, dtype=np.float64)
        num = np.asarray(num, dtype=np.int64)
        D = start.size
        if not (end.size == D and num.size == D and 1 <= D <= _lib.SARF_GRID_MAX_DIMS):
            raise ValueError("start, end and num must have the same length 1..10")
        models = {"coef": _lib.SARF_GRID_MODEL_COEF, "ref7": _lib.SARF_GRID_MODEL_REF_D7,
                  "ref10": _lib.SARF_GRID_MODEL_REF_D10}
        shrinks = {"reference": _lib.SARF_GRID_SHRINK_REFERENCE,
                   "centred": _lib.SARF_GRID_SHRINK_CENTRED}
        flags = 0
        if models[model] == _lib.SARF_GRID_MODEL_COEF:
            if ant is None or basis is None or direction is None:
                raise ValueError("model 'coef' needs ant, basis and direction")
            _need(ant, torch.float64, (Np, 3), "ant")
            _need(basis, torch.float64, (D, Np), "basis")
            per_basis = direction.dim() == 3
            _need(direction, torch.float64, (D, Np, 3) if per_basis else (Np, 3), "direction")
            flags = _lib.SARF_BATCH_DIR_PER_BASIS if per_basis else 0
        g = _lib.GridDesc()
        g.dims, g.model, g.levels, g.shrink = D, models[model], levels, shrinks[shrink]
        for d in range(D):
            g.start[d], g.end[d], g.num[d] = float(start[d]), float(end[d]), int(num[d])
        bd = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                    _lib.SARF_BP_MIXED, flags, tile, 0, *(los if los is not None else (0.0, 0.0)))
        best = np.zeros(D)
        best_e = C.c_double(0.0)
        cov = np.zeros((D, D))
        lv = (_lib.GridLevel * levels)()
        check(self.lib.sarf_grid_search(self._h, C.byref(bd), _ptr(prof), _ptr(ant), _ptr(r0),
                                        _ptr(f0), C.byref(g), _ptr(basis), _ptr(direction),
                                        best.ctypes.data_as(_P), C.byref(best_e),
                                        cov.ctypes.data_as(_P), lv, self._stream()))
        out = [dict(start=np.array(L.start[:D]), end=np.array(L.end[:D]),
                    best=np.array(L.best[:D]), entropy=float(L.entropy), gpu_ms=float(L.gpu_ms),
                    ncand=int(L.ncand)) for L in lv]
        return best, float(best_e.value), cov, out

    def entropy_grad(self, prof: torch.Tensor, ant: torch.Tensor, r0: torch.Tensor,
                     f0: torch.Tensor, grid: ImageGrid, maxWr: float, img: torch.Tensor,
                     ent: torch.Tensor, *, tile: int = 0, los: tuple[float, float] | None = None,
                     out: torch.Tensor | None = None) -> torch.Tensor:
        """dE/d(antenna positions), (Np, 3) float64, given the forward image
        `img` of the same inputs and `ent` = image_entropy(img)."""
        Np, nfft = prof.shape
        _need(prof, torch.complex64, (Np, nfft), "prof")
        _need(ant, torch.float64, (Np, 3), "ant")
        _need(r0, torch.float64, (Np,), "r0")
        _need(f0, torch.float64, (Np,), "f0")
        _need(img, torch.complex64, (grid.Nx, grid.Ny), "img")
        _need(ent, torch.float64, (3,), "ent")
        if out is None:
            out = torch.empty((Np, 3), dtype=torch.float64, device=prof.device)
        _need(out, torch.float64, (Np, 3), "out")
        d = BpDesc(grid.Nx, grid.Ny, nfft, Np, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                   _lib.SARF_BP_MIXED, 0, tile, 0, *(los if los is not None else (0.0, 0.0)))
        check(self.lib.sarf_bp_entropy_grad(self._h, C.byref(d), _ptr(prof), _ptr(ant), _ptr(r0),
                                            _ptr(f0), _ptr(img), _ptr(ent), _ptr(out),
                                            self._stream()))
        return out

    def backproject_scenes(self, scenes, *, tile: int = 0) -> list:
        """Backproject many independent scenes in one persistent launch.
        `scenes` is a sequence of (prof, ant, r0, f0, grid, maxWr[, img[, los]])
        with img None or omitted to allocate it and los an optional
        (los_x, los_y) tile hint; returns the images.  `tile`: 0 (default:
        a tile per scene when every scene has a hint along y and can take
        one, else one tile for all), one tile id for every scene, or
        _lib.SARF_SCENES_TILE_MIXED (each scene on its cheapest of tiles 15,
        16, 13)."""
        descs, keep = self._scene_descs(scenes)
        check(self.lib.sarf_bp_scenes(self._h, len(scenes), descs, tile, self._stream()))
        return keep

    def autotune_scenes(self, scenes, reps: int = 2) -> tuple[int, float]:
        """Fastest accurate batch tile for backproject_scenes on these scenes
        (sarf_autotune_scenes): returns (tile, mean ms)."""
        descs, _ = self._scene_descs(scenes)
        tile, ms = C.c_int(0), C.c_float(0.0)
        check(self.lib.sarf_autotune_scenes(self._h, len(scenes), descs, reps, C.byref(tile),
                                            C.byref(ms), self._stream()))
        return tile.value, ms.value

    def _scene_descs(self, scenes):
        descs = (_lib.SceneDesc * len(scenes))()
        keep = []
        for i, sc in enumerate(scenes):
            prof, ant, r0, f0, grid, maxWr = sc[:6]
            img = sc[6] if len(sc) > 6 and sc[6] is not None else torch.empty(
                (grid.Nx, grid.Ny), dtype=torch.complex64, device=prof.device)
            los = sc[7] if len(sc) > 7 and sc[7] is not None else (0.0, 0.0)
            Np, nfft = prof.shape
            _need(prof, torch.complex64, (Np, nfft), f"scene {i} prof")
            _need(ant, torch.float64, (Np, 3), f"scene {i} ant")
            _need(r0, torch.float64, (Np,), f"scene {i} r0")
            _need(f0, torch.float64, (Np,), f"scene {i} f0")
            _need(img, torch.complex64, (grid.Nx, grid.Ny), f"scene {i} img")
            descs[i] = _lib.SceneDesc(prof.data_ptr(), ant.data_ptr(), r0.data_ptr(),
                                      f0.data_ptr(), img.data_ptr(), grid.Nx, grid.Ny, nfft, Np,
                                      grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr, *los)
            keep.append(img)
        return descs, keep

    # ----------------------------------------------------------------- PGA
    def pga(self, img: torch.Tensor, iters: int, db_down: float = 30.0, rms_stop: float = 0.0):
        """Phase gradient autofocus of img (Nx, Ny) complex64 in place
        (PGA.m:151-232).  rms_stop > 0 ends at the first iteration whose
        estimate's rms is below it, unapplied (autofocus_cpu.cpp:432; the
        later rms entries are NaN).  Returns (af_ph (Nx,) float64, rms
        (iters,) float64) device tensors."""
        Nx, Ny = img.shape
        _need(img, torch.complex64, (Nx, Ny), "img")
        af_ph = torch.empty(Nx, dtype=torch.float64, device=img.device)
        rms = torch.empty(max(iters, 1), dtype=torch.float64, device=img.device)
        d = _lib.PgaDesc(Nx, Ny, iters, db_down, rms_stop)
        check(self.lib.sarf_pga(self._h, C.byref(d), _ptr(img), _ptr(af_ph), _ptr(rms),
                                self._stream()))
        return af_ph, rms[:iters]

    def pga_scenes(self, imgs, iters: int, db_down: float = 30.0, *, out=None,
                   rms_stop: float = 0.0):
        """PGA on many images (each (Nx, Ny) complex64, sizes may differ) in
        place, one launch per step for all of them (sarf_pga_scenes).  Returns
        a list of (af_ph, rms) device tensors.  `out` may pass the list of
        (af_ph, rms) tensors to write (for graph capture)."""
        n = len(imgs)
        descs = (_lib.PgaScene * n)()
        res = []
        for i, img in enumerate(imgs):
            Nx, Ny = img.shape
            _need(img, torch.complex64, (Nx, Ny), f"image {i}")
            if out is not None:
                af_ph, rms = out[i]
                _need(af_ph, torch.float64, (Nx,), f"image {i} af_ph")
                if rms.dtype != torch.float64 or rms.numel() < max(iters, 1):
                    raise ValueError(f"image {i} rms: float64 with >= {max(iters, 1)} values")
            else:
                af_ph = torch.empty(Nx, dtype=torch.float64, device=img.device)
                rms = torch.empty(max(iters, 1), dtype=torch.float64, device=img.device)
            descs[i] = _lib.PgaScene(img.data_ptr(), af_ph.data_ptr(), rms.data_ptr(), Nx, Ny)
            res.append((af_ph, rms))
        check(self.lib.sarf_pga_scenes(self._h, n, descs, iters, db_down, rms_stop,
                                       self._stream()))
        return [(a, r[:iters]) for a, r in res]

    # --------------------------------------------------------------- focus
    def focus_host(self, ph, ant, r0, f0, grid: ImageGrid, nfft: int, maxWr: float,
                   precision: str = "mixed", out=None):
        """Whole RC + BP on host numpy arrays (what cuda_focus_SAR_image does).
        `out` optionally gives the (Nx, Ny) complex64 image array to fill;
        page-locked `ph` / `out` (e.g. numpy views of torch tensors made with
        pin_memory=True) are copied to and from the GPU without staging."""
        import numpy as np

        ph = np.ascontiguousarray(ph, dtype=np.complex64)
        Np, K = ph.shape
        ant = np.ascontiguousarray(ant, dtype=np.float64)
        r0 = np.ascontiguousarray(r0, dtype=np.float64)
        f0 = np.ascontiguousarray(f0, dtype=np.float64)
        if out is None:
            img = np.empty((grid.Nx, grid.Ny), dtype=np.complex64)
        else:
            if (out.dtype != np.complex64 or out.shape != (grid.Nx, grid.Ny)
                    or not out.flags["C_CONTIGUOUS"]):
                raise ValueError(f"out: need a contiguous complex64 array of shape "
                                 f"{(grid.Nx, grid.Ny)}")
            img = out
        d = FocusDesc(K, Np, nfft, grid.Nx, grid.Ny, grid.x0, grid.y0, grid.Wx, grid.Wy, maxWr,
                      _lib.SARF_BP_F64 if precision == "f64" else _lib.SARF_BP_MIXED)
        check(self.lib.sarf_focus(self._h, C.byref(d), ph.ctypes.data_as(_P),
                                  ant.ctypes.data_as(_P), r0.ctypes.data_as(_P),
                                  f0.ctypes.data_as(_P), img.ctypes.data_as(_P)))
        return img

    # --------------------------------------------------------------- synth
    def synth_phase_history(self, ant: torch.Tensor, r0: torch.Tensor, f0: torch.Tensor,
                            df: float, K: int, scat: torch.Tensor, amp: torch.Tensor,
                            out: torch.Tensor | None = None, method: str = "direct") -> torch.Tensor:
        """Add the phase history of point scatterers to `out` (zeros if None).
        method "direct" sums every scatterer exactly (FP64 phase); "nufft"
        spreads them through a non-uniform FFT (~1e-6 relative, for clutter)."""
        Np = ant.shape[0]
        S = scat.shape[0]
        _need(scat, torch.float64, (S, 3), "scat")
        _need(amp, torch.complex64, (S,), "amp")
        if out is None:
            out = torch.zeros((Np, K), dtype=torch.complex64, device=ant.device)
        _need(out, torch.complex64, (Np, K), "out")
        fn = (self.lib.sarf_synth_phase_history_nufft if method == "nufft"
              else self.lib.sarf_synth_phase_history)
        check(fn(self._h, K, Np, S, _ptr(scat), _ptr(amp), _ptr(ant), _ptr(r0), _ptr(f0),
                 float(df), _ptr(out), self._stream()))
        return out
