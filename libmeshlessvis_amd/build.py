"""Build libmeshless_vis.so (HIP, gfx950) in-tree with hipcc.

    python -m libmeshlessvis_amd.build          # or build() from Python

Every library source under csrc/ is compiled for gfx950 into _build/ at the
repository root, then linked into libmeshlessvis_amd/libmeshless_vis.so; the batch
renderer csrc/vis_noninteractive.cpp and the timing harness csrc/vis_timing_test.cpp become
libmeshlessvis_amd/vis_noninteractive and libmeshlessvis_amd/vis_timing_test.  The FP32
packed-math feature is switched off for device code (the gfx950 packed FP32
instructions are kept out of every kernel).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build" / "libmeshless_vis"
LIB = PKG / "libmeshless_vis.so"
CLI = PKG / "vis_noninteractive"
PROGRAMS = {"vis_noninteractive": CLI, "vis_timing_test": PKG / "vis_timing_test"}
ARCH = os.environ.get("MESHLESS_VIS_ARCH", "gfx950")

SOURCES = ["vis_runtime.cpp", "errors.cpp", "meshless.cpp", "prep.hip", "direct_sum.hip", "fft.hip", "mfma.hip", "mfma_binned.hip"]
# fft_kernels.hip is compiled once per C2R kernel variant (pass, odd radices, largest block), each
# object holding one kernel, so that the variants compile in parallel.
FFT_VARIANTS = [(p, o, t) for p in (0, 1) for o in (0, 1) for t in (512, 1024)]
UNITS = [(name, name, []) for name in SOURCES] + [
    ("fft_kernels.hip", f"fft_kernels.{'xy'[p]}{o}_{t}.hip",
     [f"-DMV_FFT_PASS={p}", f"-DMV_FFT_ODD={o}", f"-DMV_FFT_MAXT={t}"]) for p, o, t in FFT_VARIANTS]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def _common_flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}",
        "-O3",
        "-fPIC",
        "-std=c++17",
        "-Wall",
        "-Wno-unused-result",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{ROOT / 'include'}",
    ]


def _compile(src: Path, obj: Path, defines=()) -> None:
    cmd = [_hipcc(), *_common_flags(), *defines]
    if src.suffix == ".cpp":
        cmd += ["-x", "hip"]
    cmd += ["-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    noise = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if noise:
        sys.stderr.write(f"[build] {src.name}:\n" + "\n".join(noise) + "\n")


def build(verbose: bool = True) -> Path:
    """Compile every HIP source for gfx950 and link libmeshless_vis.so; returns its path."""
    BUILD.mkdir(parents=True, exist_ok=True)
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(UNITS))) as ex:
        # The slowest units (the C2R kernels, the binned MFMA kernels) are queued first.
        for name, unit, defines in sorted(UNITS, key=lambda u: not u[0].startswith(("fft_kernels", "mfma"))):
            src = CSRC / name
            obj = BUILD / (unit + ".o")
            objs.append(obj)
            jobs.append(ex.submit(_compile, src, obj, defines))
        for j in jobs:
            j.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB}")
    build_cli(verbose)
    return LIB


def build_cli(verbose: bool = True) -> Path:
    """The programs on the C ABI - the batch renderer (vis_noninteractive.cpp) and the timing
    harness (vis_timing_test.cpp) - linked against the library beside them."""
    for name, out in PROGRAMS.items():
        tmp = out.with_suffix(".tmp")
        cmd = [_hipcc(), "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(CSRC / f"{name}.cpp"),
               f"-L{PKG}", "-lmeshless_vis", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {name}.cpp:\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
        os.replace(tmp, out)
        if verbose:
            print(f"[build] {out}")
    return CLI


if __name__ == "__main__":
    build()
