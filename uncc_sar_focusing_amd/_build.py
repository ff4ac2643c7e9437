"""Build recipe for libsarf.so (gfx950) and the host example programs.

Invoked by ``__graft_entry__.build()``; can also be run directly:
``python -m uncc_sar_focusing_amd._build``.  Everything is compiled in-tree
with ``hipcc`` so the shared library travels with the source tree.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO_DIR, "include")
LIB_PATH = os.path.join(PKG_DIR, "libsarf.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# gfx950 only; packed FP32 ops are disabled (the compiler otherwise forms
# v_pk_* FP32 instructions from scalar float code).
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        ln for ln in proc.stdout.splitlines() if "packed-fp32-ops" not in ln
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


def build_library(verbose: bool = True) -> str:
    """Compile every HIP/C++ source under csrc/ (one hipcc per source, in
    parallel) and link them into libsarf.so (always rebuilds)."""
    from concurrent.futures import ThreadPoolExecutor

    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(
        glob.glob(os.path.join(CSRC, "*.cpp"))
    )
    obj_dir = os.path.join(PKG_DIR, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    objs = [os.path.join(obj_dir, os.path.basename(s) + ".o") for s in srcs]
    jobs = max(1, min(8, len(srcs), os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", f"-I{INCLUDE}", f"-I{CSRC}",
                                       "-o", so[1], so[0]], verbose), zip(srcs, objs)))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lz"], verbose)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


PROBE_PATH = os.path.join(PKG_DIR, "libsarf_probe.so")


def build_probe(verbose: bool = True) -> str:
    """Compile the peak-rate microbenchmarks (probe/) into libsarf_probe.so."""
    src = os.path.join(PKG_DIR, "probe", "sarf_probe.hip")
    tmp = PROBE_PATH + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, src], verbose)
    os.replace(tmp, PROBE_PATH)
    return PROBE_PATH


def build_examples(verbose: bool = True) -> list[str]:
    """Compile the C++ host examples against include/ and libsarf.so."""
    out = []
    ex_dir = os.path.join(PKG_DIR, "examples")
    for src in sorted(glob.glob(os.path.join(ex_dir, "*.cpp"))):
        exe = os.path.splitext(src)[0]
        cmd = [
            HIPCC, "-O2", "-std=c++17", f"-I{INCLUDE}", "-o", exe, src,
            f"-L{PKG_DIR}", "-lsarf", f"-Wl,-rpath,{PKG_DIR}", "-Wl,-rpath,$ORIGIN/..",
        ]
        _run(cmd, verbose)
        out.append(exe)
    return out


def build_oracle(verbose: bool = True) -> str | None:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = os.path.join(REPO_DIR, "oracle")
    if not os.path.exists(os.path.join(odir, "Makefile")) or shutil.which("make") is None:
        return None
    _run(["make", "-C", odir, "-B", "--no-print-directory"], verbose)
    return os.path.join(odir, "build", "libsarf_oracle.so")


def main() -> int:
    build_library(True)
    build_probe(True)
    build_examples(True)
    build_oracle(True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
