"""Benchmark: range compression + mixed-precision backprojection throughput.

Workload (N = 1): 4096 pulses x 2048 frequency samples (f0 9.6 GHz, 640 MHz,
4 deg circular aperture at 10 km ground range / 7.5 km altitude), Nfft
16384, 2048 x 2048 image at 0.05 m, 100 seeded point targets plus clutter
on every 4th grid cell.  A step is one range compression of the resident
phase history plus one backprojection of all pulses into the image.

value = pixel*pulses per second (Gpixel*pulses/s) over the timed steps,
counted only if the image is within 1e-4 relative L2 of the all-FP64 GPU
path on the same input with identical peak pixels (reported in the line).

Also reported: roofline of the backprojection kernel (algorithmic 30 FP32
flops + 3 transcendentals per pixel*pulse, from
src/cpuBackProjection/cpuBackProjection.hpp:70-87) and the CPU baseline
(the oracle's FP32-faithful restatement of the reference CPU path on a
bounded pulse subset of the same workload, all allowed host threads).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 vector peak
PEAK_TRANS_TOPS = 9.83     # 256 CU x 16 lanes/clk x 2.4 GHz (quarter-rate transcendentals)
# Non-packed FP32 FMA rate measured by `bench.py --peaks` on MI355X (profiles/gfx950_peaks.md):
# packed FP32 code generation is disabled, so this is the attainable FP32 ceiling.
MEASURED_NONPACKED_TFLOPS = 72.68
FLOPS_PER_PP = 30.0
TRANS_PER_PP = 3.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_inputs(ctx, cfg, torch, dev):
    """Phase history on the GPU: point targets and clutter synthesised in
    FP64-phase on the device (synth kernel)."""
    from uncc_sar_focusing_amd.scenes import grid_clutter

    geom = cfg.geometry()
    grid = cfg.grid()
    tg = cfg.targets()
    t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    ant, r0, f0 = t(geom.ant), t(geom.r0), t(geom.f0)
    ph = ctx.synth_phase_history(ant, r0, f0, geom.df, geom.K, t(tg.pos),
                                 t(tg.amp.astype(np.complex64)))
    n_clutter = 0
    if cfg.clutter_stride and os.environ.get("SARF_BENCH_CLUTTER", "1") != "0":
        cl = grid_clutter(grid, cfg.clutter_stride, cfg.clutter_db, cfg.seed)
        n_clutter = cl.pos.shape[0]
        ctx.synth_phase_history(ant, r0, f0, geom.df, geom.K, t(cl.pos),
                                t(cl.amp.astype(np.complex64)), out=ph, method="nufft")
    torch.cuda.synchronize()
    return geom, grid, tg, ant, r0, f0, ph, n_clutter


def host_threads() -> int:
    """Threads the CPU legs use: the process's CPU affinity, at most 16.  A
    GPU of this machine comes with a 16-core host share (OMP_NUM_THREADS is
    16 there), so more threads would time cores another GPU's host work
    owns."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info() -> dict:
    """What the CPU baseline ran on: logical CPUs, the affinity count, the
    OMP_NUM_THREADS setting and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        aff = None
    return {"nproc": os.cpu_count(), "affinity": aff,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "cpu_model": model,
            "threads_used": host_threads(),
            "thread_cap": "16 = the host-core share of one GPU on this machine"}


def cpu_baseline(cfg, geom, grid, ph_host, target_s=4.0):
    """Oracle FP32-faithful RC + BP (reference loop order and float-isms) on
    the first P pulses of the workload, full image: all allowed threads
    (median of 3 runs, the reported value) and one thread (median of 3, on
    fewer pulses).  BP cost is linear in pulses, so rates are per
    pixel*pulse."""
    import oracle

    threads = host_threads()

    def run(P, nt):
        oracle.set_threads(nt)
        t0 = time.perf_counter()
        prof = oracle.range_compress_f32(ph_host[:P], cfg.nfft)
        oracle.bp_f32_faithful(prof, geom.ant[:P], geom.r0[:P], geom.f0[:P], grid, geom.maxWr)
        return time.perf_counter() - t0

    def sized(nt, target):
        P0 = 32 if nt > 1 else 2
        t_cal = run(P0, nt)
        return int(min(geom.Np, max(P0, round(P0 * target / max(t_cal, 1e-3)))))

    P = sized(threads, target_s)
    dts = sorted(run(P, threads) for _ in range(3))
    rate = grid.Nx * grid.Ny * P / dts[1]
    P1 = sized(1, 2.0)
    dts1 = sorted(run(P1, 1) for _ in range(3))
    rate1 = grid.Nx * grid.Ny * P1 / dts1[1]
    oracle.set_threads(threads)
    return {
        "value": rate / 1e9,
        "unit": "Gpixel*pulses/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle FP32-faithful RC+BP (reference loop order, float dR) on the first "
                   f"{P} of {geom.Np} pulses, full {grid.Nx}x{grid.Ny} image, "
                   f"{threads} OpenMP threads, median of 3 runs ({dts[1]:.1f} s each)"),
        "single_thread_value": rate1 / 1e9,
        "single_thread_sample": f"first {P1} pulses, 1 thread, median of 3 ({dts1[1]:.2f} s)",
        "host": host_info(),
    }


def measured_traffic(config: str, shape):
    """HBM bytes per bp_mixed launch from the committed FETCH_SIZE/WRITE_SIZE
    summary, if it was measured on this workload with this launch shape."""
    path = os.path.join(REPO, "profiles", f"bp_mixed_{config}_hbm_counters.md")
    if not shape or not os.path.exists(path):
        return None, None
    vals = {}
    for ln in open(path):
        if ln.startswith("traffic_bytes_per_launch[") and "]:" in ln:
            k, v = ln.split("]:", 1)
            vals[k[len("traffic_bytes_per_launch["):]] = int(v)
    want = f"{shape.get('tile')},{shape.get('pulse_split')}"
    src = f"profiles/{os.path.basename(path)}"
    if want not in vals:
        return None, f"{src} has no counters for tile,splits {want} (has {sorted(vals)})"
    return vals[want], f"{src} (tile,splits {want})"


def probe_main(args) -> int:
    """--calibrate-copy / --peaks: the microbenchmarks of libsarf_probe.so."""
    import ctypes as C
    import subprocess

    path = os.path.join(REPO, "uncc_sar_focusing_amd", "libsarf_probe.so")
    if not os.path.exists(path):
        raise ImportError(f"{path} missing: run __graft_entry__.build()")
    lib = C.CDLL(path)
    for fn in (lib.sarf_probe_fma, lib.sarf_probe_trans):
        fn.argtypes, fn.restype = [C.c_int, C.POINTER(C.c_double)], C.c_int
    lib.sarf_probe_copy.argtypes = [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    lib.sarf_probe_copy.restype = C.c_int
    v = C.c_double()
    if args.calibrate_copy:
        rc = lib.sarf_probe_copy(0, 256 << 20, 1, C.byref(v))
        log(f"bench: copy calibration rc={rc}: 2 launches of k_copy, each 256 MiB read + "
            f"256 MiB written ({v.value:.0f} GB/s)")
        return 0 if rc == 0 else 1
    out = {}
    for name, call in (("fp32_fma_tflops_nonpacked", lambda: lib.sarf_probe_fma(0, C.byref(v))),
                       ("sin_cos_tops", lambda: lib.sarf_probe_trans(0, C.byref(v))),
                       ("copy_gbs_1gib", lambda: lib.sarf_probe_copy(0, 1 << 30, 5, C.byref(v)))):
        rc = call()
        if rc != 0:
            log(f"bench: probe {name} failed ({rc})")
            return 1
        out[name] = round(v.value, 2)
        log(f"bench: probe {name} = {v.value:.2f}")
    try:
        info = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
        agent = info[info.index("gfx950"):] if "gfx950" in info else ""
        for key, label in (("Compute Unit:", "rocminfo_compute_units"),
                           ("Max Clock Freq. (MHz):", "rocminfo_max_clock_mhz")):
            for ln in agent.splitlines():
                if key in ln:
                    out[label] = int(ln.split(":")[1].split()[0])
                    break
    except (OSError, ValueError, subprocess.SubprocessError) as e:
        out["rocminfo_error"] = str(e)
    out["published"] = {"fp32_vector_tflops": PEAK_FP32_TFLOPS, "trans_tops": PEAK_TRANS_TOPS,
                        "hbm_gbs": 8000.0}
    print(json.dumps({"peaks": out}), flush=True)
    return 0


def c0_leg(ctx, torch, cpu: bool):
    """The 512-pulse x 424-sample, Nfft 4096, 256^2 @ 0.1 m scene (5 targets,
    20 dB SNR) in full: GPU RC+BP time, rel-L2 and peaks against the
    all-FP64 GPU path.  CPU leg (when `cpu`): the oracle's FP32-faithful
    restatement of the reference's CPU path timed on 1 and all host threads
    (median of 3) and its own rel-L2 against the same FP64 image (the
    reference's float-dR error)."""
    from uncc_sar_focusing_amd.scenes import CONFIGS, add_noise

    cfg = CONFIGS["C0"]
    geom, grid, tg = cfg.geometry(), cfg.grid(), cfg.targets()
    dev = "cuda:0"
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    ant, r0, f0 = t(geom.ant), t(geom.r0), t(geom.f0)
    clean = ctx.synth_phase_history(ant, r0, f0, geom.df, geom.K, t(tg.pos),
                                    t(tg.amp.astype(np.complex64)), method="direct")
    ph = add_noise(clean.cpu().numpy().astype(np.complex128), cfg.snr_db,
                   cfg.seed).astype(np.complex64)
    phd = t(ph)
    los = geom.los(grid.x0, grid.y0)
    prof = torch.empty((geom.Np, cfg.nfft), dtype=torch.complex64, device=dev)
    img = torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64, device=dev)
    for _ in range(3):
        ctx.range_compress(phd, cfg.nfft, out=prof)
        ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, out=img, los=los)
    n = 50
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(n):
        ctx.range_compress(phd, cfg.nfft, out=prof)
        ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, out=img, los=los)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n
    pp = grid.Nx * grid.Ny * geom.Np
    prof64 = ctx.range_compress(phd.to(torch.complex128), cfg.nfft)
    ref = ctx.backproject(prof64, ant, r0, f0, grid, geom.maxWr, precision="f64").cpu().numpy()
    gimg = img.cpu().numpy()

    def rel(a):
        return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))

    # range compression at Nfft = K (the reference callback's default length,
    # not a power of two: Bluestein) beside the power-of-two Nfft
    import oracle

    prof_k = ctx.range_compress(phd, geom.K)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(n):
        ctx.range_compress(phd, geom.K, out=prof_k)
    e1.record()
    torch.cuda.synchronize()
    rc_k_ms = e0.elapsed_time(e1) / n
    e0.record()
    for _ in range(n):
        ctx.range_compress(phd, cfg.nfft, out=prof)
    e1.record()
    torch.cuda.synchronize()
    rc_ms = e0.elapsed_time(e1) / n
    rc_k_err = oracle.rel_l2(prof_k.cpu().numpy(), oracle.range_compress_f64(ph, geom.K))
    # exact matched filter (O(K) per pixel*pulse) on the same data
    mimg = ctx.matched_filter(phd, ant, r0, f0, geom.df, grid)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        ctx.matched_filter(phd, ant, r0, f0, geom.df, grid, out=mimg)
    e1.record()
    torch.cuda.synchronize()
    mf_ms = e0.elapsed_time(e1) / 5
    mimg = mimg.cpu().numpy()
    out = {
        "workload": (f"C0: {geom.Np} pulses x {geom.K} samples, Nfft {cfg.nfft}, "
                     f"{grid.Nx}x{grid.Ny} @ {cfg.spacing} m, {tg.pos.shape[0]} targets, 20 dB SNR"),
        "gpu_ms_rc_bp": round(ms, 4),
        "gpu_Gpixel_pulses_per_s": round(pp / (ms * 1e-3) / 1e9, 2),
        "gpu_rel_l2_vs_fp64": rel(gimg),
        "rc_ms_nfft_eq_K": round(rc_k_ms, 4),
        "rc_ms_nfft_pow2": round(rc_ms, 4),
        "rc_nfft_eq_K_rel_l2_vs_oracle": rc_k_err,
        "gpu_peaks_match": local_peaks_np(np.abs(gimg), grid, tg.pos)
        == local_peaks_np(np.abs(ref), grid, tg.pos),
        "matched_filter_gpu_ms": round(mf_ms, 3),
        "matched_filter_Gpixel_pulse_samples_per_s": round(pp * geom.K / (mf_ms * 1e-3) / 1e9, 1),
        # BP of 1/Nfft-scaled profiles = MF / Nfft up to the interpolation loss
        "fp64_bp_x_nfft_rel_l2_vs_matched_filter": float(np.linalg.norm(ref * cfg.nfft - mimg)
                                                         / np.linalg.norm(mimg)),
    }
    if cpu:
        import oracle

        res = {}
        for nt in (1, host_threads()):
            oracle.set_threads(nt)
            dts = []
            for _ in range(3):
                t0 = time.perf_counter()
                fimg = oracle.bp_f32_faithful(oracle.range_compress_f32(ph, cfg.nfft), geom.ant,
                                              geom.r0, geom.f0, grid, geom.maxWr)
                dts.append(time.perf_counter() - t0)
            res[nt] = sorted(dts)[1]
        oracle.set_threads(host_threads())
        out.update({
            "cpu_faithful_ms_1_thread": round(res[1] * 1e3, 1),
            "cpu_faithful_ms_all_threads": round(res[host_threads()] * 1e3, 1),
            "cpu_threads": host_threads(),
            "cpu_faithful_rel_l2_vs_fp64": rel(fimg),
            "gpu_speedup_vs_cpu_all_threads": round(res[host_threads()] * 1e3 / ms, 1),
        })
    log(f"bench: C0 {out}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="C1")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--no-autofocus", action="store_true", help="skip the autofocus legs")
    ap.add_argument("--no-serving", action="store_true", help="skip the uneven-scene leg")
    ap.add_argument("--no-c0", action="store_true", help="skip the full 512-pulse C0 leg")
    ap.add_argument("--leg", choices=["c0", "autofocus", "gradient", "serving"], default=None,
                    help="run only this leg and print its JSON object (profiling runs)")
    ap.add_argument("--no-tune", action="store_true",
                    help="backproject with the default launch shape (no sarf_autotune)")
    ap.add_argument("--shape", default=None,
                    help="TILE,SPLIT: backproject with this launch shape (no tuning), e.g. to "
                         "profile exactly the shape a tuned run picked")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the probe copy kernel: 2 launches, each reading and writing "
                         "256 MiB (reference for HBM byte counters)")
    ap.add_argument("--peaks", action="store_true",
                    help="only measure FP32 FMA, v_sin/v_cos and copy rates and print them")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus {args.gpus}); "
              "sharding over several GPUs is designed but not implemented", file=sys.stderr)
        return 2

    if args.calibrate_copy or args.peaks:
        return probe_main(args)

    import torch

    from uncc_sar_focusing_amd.api import Context
    from uncc_sar_focusing_amd.scenes import CONFIGS

    if args.leg:
        # one leg alone (for profiling runs): its JSON object on stdout
        ctx = Context(0)
        legs = {"c0": lambda: c0_leg(ctx, torch, cpu=not args.no_cpu),
                "autofocus": lambda: autofocus_leg(ctx, torch),
                "gradient": lambda: gradient_leg(ctx, torch, check=not args.no_check),
                "serving": lambda: serving_leg(ctx, torch)}
        out = legs[args.leg]()
        print(json.dumps({"leg": args.leg, args.leg: out}), flush=True)
        ctx.close()
        return 0

    cfg = CONFIGS[args.config]
    dev = "cuda:0"
    ctx = Context(0)
    log(f"bench: config {cfg.name}: Np={cfg.Np} K={cfg.K} Nfft={cfg.nfft} "
        f"image {cfg.n_pix}^2 @ {cfg.spacing} m")
    geom, grid, tg, ant, r0, f0, ph, n_clutter = make_inputs(ctx, cfg, torch, dev)
    log(f"bench: inputs ready ({tg.pos.shape[0]} targets, {n_clutter} clutter scatterers)")

    los = geom.los(grid.x0, grid.y0)
    prof = torch.empty((geom.Np, cfg.nfft), dtype=torch.complex64, device=dev)
    img = torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64, device=dev)

    shape = {}
    if args.shape:
        t_, s_ = (int(v) for v in args.shape.split(","))
        shape = {"tile": t_, "pulse_split": s_}

    def step():
        ctx.range_compress(ph, cfg.nfft, out=prof)
        ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, out=img, los=los, **shape)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    tune = None
    if shape:
        tune = dict(shape, source="--shape")
    elif not args.no_tune:
        # launch-shape tuning on the resident inputs (outside the timed
        # region); later default-shaped calls on this geometry use the pick
        tune, _ = ctx.autotune(prof, ant, r0, f0, grid, geom.maxWr, los=los)
        tune = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tune.items()}
        log(f"bench: autotune {tune}")
        step()
        torch.cuda.synchronize()

    # timed region: K steps, synchronised on both sides
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev[0].record()
    for i in range(args.steps):
        ctx.range_compress(ph, cfg.nfft, out=prof)
        ev[2 * i + 1].record()
        ctx.backproject(prof, ant, r0, f0, grid, geom.maxWr, out=img, los=los, **shape)
        ev[2 * i + 2].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_rc = sum(ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)) / args.steps
    t_bp = sum(ev[2 * i + 1].elapsed_time(ev[2 * i + 2]) for i in range(args.steps)) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    pp = grid.Nx * grid.Ny * geom.Np
    value = pp / (ms_per_step * 1e-3) / 1e9
    log(f"bench: {ms_per_step:.3f} ms/step (rc {t_rc:.3f} ms, bp {t_bp:.3f} ms) -> "
        f"{value:.1f} Gpixel*pulses/s")

    # accuracy gate against the all-FP64 GPU path on the same input
    rel = None
    peaks_match = None
    c1_oracle = None
    if not args.no_check:
        img32 = img.clone()
        prof64 = ctx.range_compress(ph.to(torch.complex128), cfg.nfft)
        img64 = ctx.backproject(prof64, ant, r0, f0, grid, geom.maxWr, precision="f64")
        torch.cuda.synchronize()
        diff = (img32.to(torch.complex128) - img64)
        rel = float(torch.linalg.vector_norm(diff) / torch.linalg.vector_norm(img64))
        a32 = img32.abs().cpu().numpy()
        a64 = img64.abs().cpu().numpy()
        peaks_match = bool(local_peaks_np(a32, grid, tg.pos) == local_peaks_np(a64, grid, tg.pos)
                           and np.argmax(a32) == np.argmax(a64))
        del prof64, img64
        log(f"bench: rel-L2 vs FP64 GPU path {rel:.3e}, peaks match {peaks_match}")
        c1_oracle = oracle_subset_check(ctx, torch, ph, prof, ant, r0, f0, grid, geom.maxWr,
                                        cfg.nfft, los)
        log(f"bench: 128-pulse subset vs CPU FP64 oracle {c1_oracle}")

    # host-buffer path (sarf_focus, what cuda_focus_SAR_image does): includes
    # the PCIe copies of the phase history in and the image out; not `value`
    ph_host = ph.cpu().numpy()
    ctx.focus_host(ph_host, geom.ant, geom.r0, geom.f0, grid, cfg.nfft, geom.maxWr)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.focus_host(ph_host, geom.ant, geom.r0, geom.f0, grid, cfg.nfft, geom.maxWr)
    host_ms = (time.perf_counter() - t0) * 1e3 / 3
    log(f"bench: host-buffer path (sarf_focus, PCIe included) {host_ms:.2f} ms -> "
        f"{pp / (host_ms * 1e-3) / 1e9:.1f} Gpixel*pulses/s")
    # the same with page-locked caller buffers (no staging copies)
    ph_pin = torch.empty(tuple(ph.shape), dtype=torch.complex64, pin_memory=True)
    ph_pin.copy_(ph)
    img_pin = torch.empty((grid.Nx, grid.Ny), dtype=torch.complex64, pin_memory=True)
    ph_pin_np, img_pin_np = ph_pin.numpy(), img_pin.numpy()
    ctx.focus_host(ph_pin_np, geom.ant, geom.r0, geom.f0, grid, cfg.nfft, geom.maxWr,
                   out=img_pin_np)
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.focus_host(ph_pin_np, geom.ant, geom.r0, geom.f0, grid, cfg.nfft, geom.maxWr,
                       out=img_pin_np)
    host_pin_ms = (time.perf_counter() - t0) * 1e3 / 3
    del ph_pin, img_pin, ph_pin_np, img_pin_np
    log(f"bench: host-buffer path, page-locked buffers {host_pin_ms:.2f} ms -> "
        f"{pp / (host_pin_ms * 1e-3) / 1e9:.1f} Gpixel*pulses/s")

    # roofline of the backprojection kernel
    achieved = FLOPS_PER_PP * pp / (t_bp * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic(cfg.name, tune)
    roof = {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "kernel": "bp_mixed",
        "trans_frac": round(TRANS_PER_PP * pp / (t_bp * 1e-3) / 1e12 / PEAK_TRANS_TOPS, 4),
        "nonpacked_fp32_measured": MEASURED_NONPACKED_TFLOPS,
        "frac_of_nonpacked": round(achieved / MEASURED_NONPACKED_TFLOPS, 4),
    }

    autofocus = None
    gradient = None
    if not args.no_autofocus:
        autofocus = autofocus_leg(ctx, torch)
        gradient = gradient_leg(ctx, torch, check=not args.no_check)
    serving = None
    if not args.no_serving:
        serving = serving_leg(ctx, torch)

    c0 = None
    if not args.no_c0:
        c0 = c0_leg(ctx, torch, cpu=not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        ph_host = ph.cpu().numpy()
        cpu = cpu_baseline(cfg, geom, grid, ph_host)
        log(f"bench: cpu baseline {cpu['value']*1e3:.1f} Mpixel*pulses/s ({cpu['sample']}); "
            f"1 thread {cpu['single_thread_value']*1e3:.1f} Mpixel*pulses/s")

    line = {
        "metric": "BP Gpixel·pulses/s at ≤1e-4 rel-L2 vs FP64; autofocus time-to-focus",
        "value": round(value, 3),
        "unit": "Gpixel*pulses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 per tile-pulse geometry)",
        "data": "synthetic",
        "config": {
            "workload": (f"{cfg.name}: RC+BP, {geom.Np} pulses x {geom.K} samples, Nfft {cfg.nfft}, "
                         f"{grid.Nx}x{grid.Ny} image @ {cfg.spacing} m, {tg.pos.shape[0]} targets "
                         f"+ {n_clutter} clutter scatterers"),
            "pixel_pulses": pp,
            "rc_ms": round(t_rc, 4),
            "bp_ms": round(t_bp, 4),
            "rel_l2_vs_fp64": rel,
            "peaks_match": peaks_match,
            "cpu_oracle_128_pulses": c1_oracle,
            "launch_shape": tune,
            "host_buffer_path_ms": round(host_ms, 3),
            "host_buffer_path_Gpixel_pulses_per_s": round(pp / (host_ms * 1e-3) / 1e9, 1),
            "host_buffer_path_page_locked_ms": round(host_pin_ms, 3),
            "host_buffer_path_page_locked_Gpixel_pulses_per_s":
                round(pp / (host_pin_ms * 1e-3) / 1e9, 1),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "c0": c0,
        "autofocus": autofocus,
        "gradient_autofocus": gradient,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    ctx.close()
    return 0


def autofocus_leg(ctx, torch, levels: int = 5):
    """Grid-search autofocus on the 1024 x 1024, 512^2 scene with a seeded
    line-of-sight polynomial error a t^2 + b t^3 + c t^4.

    * Single level, the error model's own bases (t^2, t^3, t^4): all 11^3 =
      1331 candidate images in one persistent launch, scored by entropy.
    * Multi-resolution (sarf_grid_search, the level loop of the reference's
      driver): `levels` boxes of 11^3, each half the span of the previous
      around its minimiser, over the detrended bases (each minus its
      least-squares line in t): a LOS offset constant or linear in slow time
      only shifts the image, and t^3 carries a linear part, so with the raw
      bases the search also moves the scene across the image border (the
      minimum found that way is a shifted image, not a focused one).
    * Gradient autofocus on the same scene, until its entropy reaches the
      multi-resolution minimum: Adam over the same 3 coefficients, and over
      1024 per-pulse LOS offsets."""
    from uncc_sar_focusing_amd import autofocus as af
    from uncc_sar_focusing_amd.autograd import BpOperator, gradient_autofocus
    from uncc_sar_focusing_amd.workloads import los_error_scene

    sc = los_error_scene(ctx, "C2")
    g, grid = sc.geom, sc.grid
    los = g.los(grid.x0, grid.y0)
    prof = ctx.range_compress(sc.ph, sc.cfg.nfft)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    B = af.poly_basis(g.Np)
    Bd = af.detrended_basis(B)
    dir_h = af.los_directions(g.ant, grid.x0, grid.y0)
    direction = t(dir_h)
    coefs = af.coef_grid(-0.3, 0.3, 11)
    # warm-up (scratch allocation, code load) on a small candidate set
    af.grid_search(ctx, prof, sc.ant, sc.r0, sc.f0, grid, g.maxWr, t(B), direction, coefs[:16],
                   los=los)
    res = af.grid_search(ctx, prof, sc.ant, sc.r0, sc.f0, grid, g.maxWr, t(B), direction, coefs,
                         los=los)

    def image_at(coef, basis):
        ant = t(g.ant + (np.asarray(coef) @ basis)[:, None] * dir_h)
        return ctx.backproject(prof, ant, sc.r0, sc.f0, grid, g.maxWr, los=los)

    def entropy_at(coef, basis):
        return float(ctx.image_entropy(image_at(coef, basis)).cpu()[0])

    pp = grid.Nx * grid.Ny * g.Np * res.ncand
    out = {
        "workload": (f"C2: {g.Np} pulses x {g.K} samples, Nfft {sc.cfg.nfft}, {grid.Nx}^2 @ "
                     f"{sc.cfg.spacing} m, {sc.targets.pos.shape[0]} targets + {sc.n_clutter} "
                     f"clutter; LOS error a t^2 + b t^3 + c t^4"),
        "true_coef": [round(float(c), 4) for c in sc.true_coef],
        "grid_candidates": res.ncand,
        "grid_gpu_ms": round(res.gpu_ms, 3),
        "grid_Gpixel_pulses_per_s": round(pp / (res.gpu_ms * 1e-3) / 1e9, 1),
        "raw_basis_best_coef": [round(float(c), 4) for c in res.best_coef],
        "raw_basis_E_best": res.best_entropy,
        "E_nominal": entropy_at(np.zeros(3), B),
        "E_true": entropy_at(sc.true_coef, B),
    }
    # image shift of the raw-basis minimum against the focused image
    a0 = image_at(sc.true_coef, B).abs().cpu().numpy()
    a1 = image_at(res.best_coef, B).abs().cpu().numpy()
    xc = np.fft.ifft2(np.fft.fft2(a1) * np.conj(np.fft.fft2(a0))).real
    sh = np.unravel_index(int(np.argmax(xc)), xc.shape)
    out["raw_basis_best_shift_px_xy"] = [int(v if v < n // 2 else v - n)
                                         for v, n in zip(sh, xc.shape)]

    # multi-resolution search over the detrended bases
    tBd = t(Bd)
    mr = af.grid_search_multires(ctx, prof, sc.r0, sc.f0, grid, g.maxWr, [-0.3] * 3, [0.3] * 3,
                                 [11] * 3, levels, ant=sc.ant, basis=tBd, direction=direction,
                                 los=los)
    last = mr.levels[-1]
    step = (last["end"] - last["start"]) / 10.0
    E_true_d = entropy_at(sc.true_coef, Bd)
    out["multires"] = {
        "basis": "t^2, t^3, t^4 minus their least-squares line in t",
        "levels": [{"best": [round(float(v), 5) for v in L["best"]],
                    "E": round(L["entropy"], 6), "gpu_ms": round(L["gpu_ms"], 3),
                    "step": [round(float(v), 6) for v in (L["end"] - L["start"]) / 10.0]}
                   for L in mr.levels],
        "best_coef": [round(float(v), 5) for v in mr.best],
        "abs_err_vs_true": [round(float(v), 5) for v in np.abs(mr.best - sc.true_coef)],
        "final_step": [round(float(v), 6) for v in step],
        "within_one_final_step": bool(np.all(np.abs(mr.best - sc.true_coef) <= step + 1e-12)),
        "E_found": mr.best_entropy,
        "E_true": E_true_d,
        "covariance_diag": [float(v) for v in np.diag(mr.covariance)],
        "gpu_ms_total": round(mr.gpu_ms, 3),
    }
    # image precision of the backprojection engine on this configuration:
    # the focused (true-trajectory) image, mixed vs all-FP64 GPU path, and
    # against the CPU FP64 oracle on a 128-pulse subset
    ant_true = t(sc.true_geom.ant)
    img32 = ctx.backproject(prof, ant_true, sc.r0, sc.f0, grid, g.maxWr, los=los)
    prof64 = ctx.range_compress(sc.ph.to(torch.complex128), sc.cfg.nfft)
    img64 = ctx.backproject(prof64, ant_true, sc.r0, sc.f0, grid, g.maxWr, precision="f64")
    out["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(img32.to(torch.complex128) - img64)
                                  / torch.linalg.vector_norm(img64))
    a32, a64 = img32.abs().cpu().numpy(), img64.abs().cpu().numpy()
    out["peaks_match"] = bool(np.argmax(a32) == np.argmax(a64) and
                              local_peaks_np(a32, grid, sc.targets.pos)
                              == local_peaks_np(a64, grid, sc.targets.pos))
    out["cpu_oracle_128_pulses"] = oracle_subset_check(ctx, torch, sc.ph, prof, ant_true, sc.r0,
                                                       sc.f0, grid, g.maxWr, sc.cfg.nfft, los)
    del prof64, img64
    log(f"bench: grid search {res.ncand} candidates in {res.gpu_ms:.1f} ms GPU "
        f"({out['grid_Gpixel_pulses_per_s']} Gpixel*pulses/s); raw-basis best "
        f"{out['raw_basis_best_coef']} (image shift {out['raw_basis_best_shift_px_xy']} px) true "
        f"{out['true_coef']}; E best {res.best_entropy:.4f} nominal {out['E_nominal']:.4f} "
        f"true {out['E_true']:.4f}")
    log(f"bench: multires {levels} levels in {mr.gpu_ms:.1f} ms GPU: best "
        f"{out['multires']['best_coef']} |err| {out['multires']['abs_err_vs_true']} final step "
        f"{out['multires']['final_step']} (within one step: {out['multires']['within_one_final_step']}); E found {mr.best_entropy:.5f} true {E_true_d:.5f}")

    # gradient autofocus on the same scene until E <= the multi-resolution minimum
    op = BpOperator(ctx, prof, sc.r0, sc.f0, grid, g.maxWr, los)
    ant0 = sc.ant
    E_goal = mr.best_entropy
    grid_ms = mr.gpu_ms

    def first_hit(gr):
        hit = np.nonzero(gr.entropies <= E_goal)[0]
        # entropies[i] is E before step i; the time to have evaluated it is
        # bounded by the end of step i
        if not hit.size:
            return None, None
        i = int(hit[0])
        return i, float(gr.gpu_ms[i])

    make3 = lambda c: ant0 + (c @ tBd)[:, None] * direction  # noqa: E731
    c0 = torch.zeros(3, dtype=torch.float64, device="cuda:0")
    gradient_autofocus(op, make3, c0, 2, 0.01)  # warm-up
    gr3 = gradient_autofocus(op, make3, c0, 150, 0.01)
    i3, t3 = first_hit(gr3)
    makep = lambda d: ant0 + d[:, None] * direction  # noqa: E731
    p0 = torch.zeros(g.Np, dtype=torch.float64, device="cuda:0")
    grp = gradient_autofocus(op, makep, p0, 40, 0.003)
    ip, tp = first_hit(grp)
    out.update({
        "gradient_3coef": {
            "params": "the same 3 detrended-basis coefficients, Adam lr 1e-2, 150 steps",
            "iter_to_E_grid": i3,
            "gpu_ms_to_E_grid": None if t3 is None else round(t3, 3),
            "over_grid_time": None if t3 is None else round(t3 / grid_ms, 4),
            "ms_per_iter": round(float(gr3.gpu_ms[-1]) / gr3.iters, 3),
            "final_coef": [round(float(v), 5) for v in gr3.params.cpu().numpy()],
            "E_final": float(gr3.entropies[-1]),
        },
        "gradient_per_pulse": {
            "params": "per-pulse LOS offset (1024 unknowns), Adam lr 3e-3, 40 steps",
            "iter_to_E_grid": ip,
            "gpu_ms_to_E_grid": None if tp is None else round(tp, 3),
            "over_grid_time": None if tp is None else round(tp / grid_ms, 4),
            "ms_per_iter": round(float(grp.gpu_ms[-1]) / grp.iters, 3),
            "E_final": float(grp.entropies[-1]),
        },
    })
    log(f"bench: gradient autofocus to E_grid: 3 coefficients at step {i3} ({t3} ms GPU, "
        f"final {out['gradient_3coef']['final_coef']}); per-pulse at step {ip} ({tp} ms GPU); "
        f"grid {grid_ms:.1f} ms")
    return out


def _lib_mixed_tile() -> int:
    from uncc_sar_focusing_amd import _lib

    return _lib.SARF_SCENES_TILE_MIXED


def oracle_subset_check(ctx, torch, ph, prof, ant, r0, f0, grid, maxWr, nfft, los,
                        npulses: int = 128, first: int = 0):
    """The GPU image (range compression + mixed-precision backprojection) of
    pulses first..first+npulses-1 against the CPU FP64 oracle's range
    compression + backprojection of the same phase history: rel-L2 and
    brightest-pixel equality.  Backprojection is linear in the pulses, so a
    contiguous subset checks the full-size image geometry at CPU cost."""
    import oracle

    sl = slice(first, first + npulses)
    p = prof[sl].contiguous()  # per-pulse transforms: equal to compressing the subset alone
    a, r, f = ant[sl].contiguous(), r0[sl].contiguous(), f0[sl].contiguous()
    img = ctx.backproject(p, a, r, f, grid, maxWr, los=los).cpu().numpy()
    t0 = time.perf_counter()
    ref = oracle.bp_f64(oracle.range_compress_f64(ph[sl].cpu().numpy(), nfft), a.cpu().numpy(),
                        r.cpu().numpy(), f.cpu().numpy(), grid, maxWr)
    dt = time.perf_counter() - t0
    return {"pulses": f"{first}..{first + npulses - 1}",
            "rel_l2": oracle.rel_l2(img, ref),
            "argmax_match": bool(np.argmax(np.abs(img)) == np.argmax(np.abs(ref))),
            "oracle_s": round(dt, 2)}


def serving_leg(ctx, torch, n_scenes: int = 64, pga_iters: int = 5, n_streams: int = 4):
    """64 uneven scenes (pulses 256-4096, samples 256-2048, sides 128-1024)
    with an 8th-order azimuth phase error, each range-compressed,
    backprojected and corrected by 5 PGA iterations.  Batched: all scenes'
    range compression in one sarf_range_compress_scenes call (one launch
    per transform length), their backprojection in one persistent launch,
    then all scenes' PGA in one
    sarf_pga_scenes call (each PGA step one launch over all images); also
    timed: the same batch with per-image PGA calls spread over `n_streams`
    streams (one context each); latency: each scene alone on one stream, in
    steady state (after a first pass over all scenes)."""
    from uncc_sar_focusing_amd.api import Context
    from uncc_sar_focusing_amd.workloads import uneven_scenes

    scenes = uneven_scenes(ctx, n_scenes)
    pctx = [Context(0) for _ in range(n_streams)]
    streams = [torch.cuda.Stream() for _ in range(n_streams)]
    total_pp = sum(s.pixel_pulses for s in scenes)
    profs = [torch.empty((s.geom.Np, s.nfft), dtype=torch.complex64, device="cuda:0")
             for s in scenes]
    imgs = [torch.empty((s.grid.Nx, s.grid.Ny), dtype=torch.complex64, device="cuda:0")
            for s in scenes]
    pga_out = [(torch.empty(s.grid.Nx, dtype=torch.float64, device="cuda:0"),
                torch.empty(pga_iters, dtype=torch.float64, device="cuda:0")) for s in scenes]
    descs = [(profs[i], s.ant, s.r0, s.f0, s.grid, s.geom.maxWr, imgs[i],
              s.geom.los(s.grid.x0, s.grid.y0)) for i, s in enumerate(scenes)]

    phs = [s.ph for s in scenes]
    nffts = [s.nfft for s in scenes]
    # batch tile for the one-launch backprojection of all scenes
    ctx.range_compress_scenes(phs, nffts, out=profs)
    scene_tile, _ = ctx.autotune_scenes(descs)

    def rc_bp(e_rc=None):
        ctx.range_compress_scenes(phs, nffts, out=profs)
        if e_rc is not None:
            e_rc.record()
        ctx.backproject_scenes(descs, tile=scene_tile)

    def batch(timed=False):
        e_rc = torch.cuda.Event(enable_timing=True) if timed else None
        rc_bp(e_rc)
        e_bp = torch.cuda.Event(enable_timing=True)
        e_bp.record()
        ctx.pga_scenes(imgs, pga_iters, out=pga_out)
        return e_rc, e_bp

    order = sorted(range(len(imgs)), key=lambda i: -imgs[i].numel())  # big scenes first

    def batch_streams():
        rc_bp()
        e_bp = torch.cuda.Event()
        e_bp.record()
        for st in streams:
            st.wait_event(e_bp)
        for j, i in enumerate(order):
            with torch.cuda.stream(streams[j % n_streams]):
                pctx[j % n_streams].pga(imgs[i], pga_iters)
        cur = torch.cuda.current_stream()
        for st in streams:
            ev = torch.cuda.Event()
            ev.record(st)
            cur.wait_event(ev)

    # warm-up: scratch buffers, tables, the uploaded scene lists, the per-image
    # contexts.  The batch goes last: the timed batch then starts at the clocks
    # of continuous load (after lighter work or idle the GPU runs the batch's
    # backprojection up to 10 % slower while its clocks ramp up)
    batch_streams()
    batch()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    e_rc, e_bp = batch(timed=True)
    e1.record()
    e1.synchronize()
    t_batch = e0.elapsed_time(e1)
    t_rcbp = e0.elapsed_time(e_bp)
    t_rc = e0.elapsed_time(e_rc)
    e0.record()
    batch_streams()
    e1.record()
    e1.synchronize()
    t_streams = e0.elapsed_time(e1)
    def one_scene(i, s):
        p = ctx.range_compress(s.ph, s.nfft, out=profs[i])
        im = ctx.backproject(p, s.ant, s.r0, s.f0, s.grid, s.geom.maxWr, out=imgs[i],
                             los=s.geom.los())
        ctx.pga(im, pga_iters)

    # steady state: a first pass sizes the context's scratch (the first call
    # of a larger scene allocates); its maximum is reported beside it
    first = []
    for i, s in enumerate(scenes):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        one_scene(i, s)
        b.record()
        b.synchronize()
        first.append(a.elapsed_time(b))
    lat = []
    for i, s in enumerate(scenes):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        one_scene(i, s)
        b.record()
        b.synchronize()
        lat.append(a.elapsed_time(b))
    lat = np.array(lat)
    # the same batch captured once into a HIP graph and replayed
    t_graph = None
    try:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            batch()
        graph.replay()
        torch.cuda.synchronize()
        e0.record()
        graph.replay()
        e1.record()
        e1.synchronize()
        t_graph = e0.elapsed_time(e1)
    except RuntimeError as e:  # capture unsupported here: report, keep the eager numbers
        log(f"bench: serving graph capture failed: {e}")
    # precision of the batched backprojection (before PGA) against the
    # all-FP64 path, every scene
    ctx.range_compress_scenes(phs, nffts, out=profs)
    ctx.backproject_scenes(descs, tile=scene_tile)
    rels, peaks = [], []
    for i, s in enumerate(scenes):
        p64 = ctx.range_compress(s.ph.to(torch.complex128), s.nfft)
        i64 = ctx.backproject(p64, s.ant, s.r0, s.f0, s.grid, s.geom.maxWr, precision="f64")
        rels.append(float(torch.linalg.vector_norm(imgs[i].to(torch.complex128) - i64)
                          / torch.linalg.vector_norm(i64)))
        peaks.append(int(torch.argmax(imgs[i].abs())) == int(torch.argmax(i64.abs())))
        del p64, i64
    big = max(range(len(scenes)), key=lambda i: scenes[i].pixel_pulses)
    sb = scenes[big]
    big_oracle = oracle_subset_check(ctx, torch, sb.ph, profs[big], sb.ant, sb.r0, sb.f0, sb.grid,
                                     sb.geom.maxWr, sb.nfft, sb.geom.los(sb.grid.x0, sb.grid.y0))
    big_oracle["scene"] = big
    # focus check on the first scenes (outside the timed loop)
    ent_before, ent_after = [], []
    for i, s in enumerate(scenes[:8]):
        p = ctx.range_compress(s.ph, s.nfft, out=profs[i])
        im = ctx.backproject(p, s.ant, s.r0, s.f0, s.grid, s.geom.maxWr, out=imgs[i])
        ent_before.append(float(ctx.image_entropy(im).cpu()[0]))
        ctx.pga(im, pga_iters)
        ent_after.append(float(ctx.image_entropy(im).cpu()[0]))
    out = {
        "workload": (f"C4: {n_scenes} scenes, pulses {min(s.geom.Np for s in scenes)}-"
                     f"{max(s.geom.Np for s in scenes)}, samples {min(s.geom.K for s in scenes)}-"
                     f"{max(s.geom.K for s in scenes)}, sides {min(s.grid.Nx for s in scenes)}-"
                     f"{max(s.grid.Nx for s in scenes)}, 8th-order azimuth phase error, "
                     f"{pga_iters} PGA iterations"),
        "pixel_pulses": int(total_pp),
        "batch_gpu_ms": round(t_batch, 3),
        "batch_rc_bp_gpu_ms": round(t_rcbp, 3),
        "batch_rc_gpu_ms": round(t_rc, 3),
        "rc": "batched (sarf_range_compress_scenes)",
        "batch_Gpixel_pulses_per_s": round(total_pp / (t_batch * 1e-3) / 1e9, 1),
        "batch_rc_bp_Gpixel_pulses_per_s": round(total_pp / (t_rcbp * 1e-3) / 1e9, 1),
        "scenes_tile": ("mixed: each scene on its cheapest of tiles 15, 16, 13"
                        if scene_tile == _lib_mixed_tile() else scene_tile),
        "pga": "batched (sarf_pga_scenes)",
        "batch_per_image_pga_streams_gpu_ms": round(t_streams, 3),
        "pga_streams": n_streams,
        "batch_graph_replay_gpu_ms": None if t_graph is None else round(t_graph, 3),
        "bp_rel_l2_vs_fp64_max": max(rels),
        "bp_rel_l2_vs_fp64_median": float(np.median(rels)),
        "bp_argmax_match_fp64_all_scenes": bool(all(peaks)),
        "largest_scene_cpu_oracle_128_pulses": big_oracle,
        "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "latency_first_call_ms_max": round(float(max(first)), 3),
        "entropy_before_pga_first8": [round(v, 4) for v in ent_before],
        "entropy_after_pga_first8": [round(v, 4) for v in ent_after],
    }
    log(f"bench: serving {n_scenes} scenes in {t_batch:.1f} ms GPU "
        f"({out['batch_Gpixel_pulses_per_s']} Gpixel*pulses/s incl. PGA, RC+BP "
        f"{out['batch_rc_bp_Gpixel_pulses_per_s']}); graph replay {out['batch_graph_replay_gpu_ms']} ms; "
        f"latency p50 {out['latency_ms_p50']} ms p99 {out['latency_ms_p99']} ms; "
        f"BP rel-L2 vs FP64 max {out['bp_rel_l2_vs_fp64_max']:.2e}")
    for c in pctx:
        c.close()
    return out


def gradient_leg(ctx, torch, check: bool = True):
    """Per-pulse 3-D errors (smoothed random walk, 0.05 m per axis) on the
    grid-search scene: 3072 unknowns, Adam for 200 steps through the adjoint.
    Gradient checks: central differences (h = 1e-4 m) of the FP64 GPU forward
    model on 32 random pulses, and the FP64 CPU adjoint of the oracle."""
    from uncc_sar_focusing_amd.autograd import BpOperator, gradient_autofocus
    from uncc_sar_focusing_amd.workloads import random_walk_scene

    sc = random_walk_scene(ctx, "C2")
    g, grid = sc.geom, sc.grid
    los = g.los(grid.x0, grid.y0)
    prof = ctx.range_compress(sc.ph, sc.cfg.nfft)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")  # noqa: E731
    op = BpOperator(ctx, prof, sc.r0, sc.f0, grid, g.maxWr, los)

    def E_at(ant):
        return float(ctx.image_entropy(op.image(ant)).cpu()[0])

    from uncc_sar_focusing_amd import autofocus as af
    from uncc_sar_focusing_amd.autograd import los_residual_rms, smoothing_operator

    p0 = torch.zeros((g.Np, 3), dtype=torch.float64, device="cuda:0")
    make = lambda d: sc.ant + d  # noqa: E731
    gradient_autofocus(op, make, p0, 2, 0.003)  # warm-up
    gr = gradient_autofocus(op, make, p0, 200, 0.003)
    # the same 3072 unknowns q, descended through a Gaussian smoothing of 16
    # pulses (p = S q): updates stay in the band of a smooth trajectory error
    smooth = smoothing_operator(g.Np, 16.0, "cuda:0")
    makes = lambda q: sc.ant + smooth(q)  # noqa: E731
    grs = gradient_autofocus(op, makes, p0, 200, 0.003)
    err = sc.true_geom.ant - g.ant
    d = af.los_directions(g.ant, grid.x0, grid.y0)
    est = gr.params.cpu().numpy()
    est_s = smooth(grs.params).cpu().numpy()
    out = {
        "workload": (f"C3: C2 scene ({sc.n_clutter} clutter), per-pulse 3-D errors, smoothed "
                     f"random walk sigma 0.05 m/axis, {3 * g.Np} unknowns, Adam 200 steps"),
        "E_nominal": float(gr.entropies[0]),
        "E_true": E_at(t(sc.true_geom.ant)),
        "E_final": E_at(make(gr.params)),
        "gpu_ms_total": round(float(gr.gpu_ms[-1]), 3),
        "ms_per_iter": round(float(gr.gpu_ms[-1]) / gr.iters, 3),
        "rms_err_before_m": float(np.sqrt(np.mean(err ** 2))),
        "rms_err_after_m": float(np.sqrt(np.mean((est - err) ** 2))),
        # the observable part: LOS component minus its line in slow time
        "los_detrended_rms_before_m": los_residual_rms(err, d),
        "los_detrended_rms_after_m": los_residual_rms(err - est, d),
        "smoothed": {
            "params": "3072 unknowns q, positions a + S q (Gaussian smoothing, 16 pulses)",
            "E_final": E_at(makes(grs.params)),
            "gpu_ms_total": round(float(grs.gpu_ms[-1]), 3),
            "los_detrended_rms_after_m": los_residual_rms(err - est_s, d),
        },
    }
    if check:
        import oracle

        ant = g.ant
        img = op.image(sc.ant)
        gpu_grad = ctx.entropy_grad(prof, sc.ant, sc.r0, sc.f0, grid, g.maxWr, img,
                                    ctx.image_entropy(img), los=los).cpu().numpy()
        prof64 = ctx.range_compress(sc.ph.to(torch.complex128), sc.cfg.nfft)

        def E64(a):
            im = ctx.backproject(prof64, t(a), sc.r0, sc.f0, grid, g.maxWr, precision="f64")
            return float(ctx.image_entropy(im).cpu()[0])

        rng = np.random.default_rng(0)
        h = 1e-4
        dev = []
        for n in rng.choice(g.Np, 32, replace=False):
            for ax in range(3):
                ap, am = ant.copy(), ant.copy()
                ap[n, ax] += h
                am[n, ax] -= h
                dev.append(abs((E64(ap) - E64(am)) / (2 * h) - gpu_grad[n, ax]))
        gmax = float(np.abs(gpu_grad).max())
        prof_h = prof64.cpu().numpy()
        img_ref = oracle.bp_f64(prof_h, ant, g.r0, g.f0, grid, g.maxWr)
        cpu_grad = oracle.bp_adjoint_f64(prof_h, ant, g.r0, g.f0, grid, g.maxWr, img_ref)
        out.update({
            "fd_max_abs_dev_over_max_grad": max(dev) / gmax,
            "cpu_fp64_grad_rel_l2": float(np.linalg.norm(gpu_grad - cpu_grad)
                                          / np.linalg.norm(cpu_grad)),
        })
    log(f"bench: gradient autofocus 200 steps in {out['gpu_ms_total']} ms GPU; E "
        f"{out['E_nominal']:.4f} -> {out['E_final']:.4f} (true trajectory {out['E_true']:.4f}); "
        f"LOS detrended rms {out['los_detrended_rms_before_m']:.4f} -> "
        f"{out['los_detrended_rms_after_m']:.4f} m (smoothed: E {out['smoothed']['E_final']:.4f}, "
        f"{out['smoothed']['los_detrended_rms_after_m']:.4f} m); checks "
        f"{out.get('fd_max_abs_dev_over_max_grad')}, {out.get('cpu_fp64_grad_rel_l2')}")
    return out


def local_peaks_np(absimg, grid, pos, radius=0.5):
    x, y = grid.x(), grid.y()
    out = []
    for p in pos:
        ix = np.nonzero(np.abs(x - p[0]) <= radius)[0]
        iy = np.nonzero(np.abs(y - p[1]) <= radius)[0]
        if ix.size == 0 or iy.size == 0:
            continue
        sub = absimg[np.ix_(ix, iy)]
        a, b = np.unravel_index(int(np.argmax(sub)), sub.shape)
        out.append((int(ix[a]), int(iy[b])))
    return out


if __name__ == "__main__":
    sys.exit(main())
