"""Compile the HIP library ``seissol_amd/_lib/libseissol_amd.so`` for gfx950.

Every source under ``seissol_amd/csrc`` is compiled by ``hipcc`` into an object
under ``seissol_amd/_build`` (in parallel), then linked into one shared
library.  Packed FP32 instructions are disabled for all device code.  The
proxy executable ``seissol_amd/_lib/seissol_amd_proxy`` (host code in
``seissol_amd/host``) is linked against that library.

Usage: ``python -m seissol_amd.build [-j N]``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIBDIR = PKG / "_lib"
LIBNAME = "libseissol_amd.so"
PROXY = "seissol_amd_proxy"
HOST = PKG / "host"
ARCH = os.environ.get("SDG_OFFLOAD_ARCH", "gfx950")

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", str(PKG.parent / "include"),
]


def _hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else (shutil.which("hipcc") or "hipcc")


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    cmd = [_hipcc(), *DEVICE_FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [_hipcc(), "-x", "hip", *DEVICE_FLAGS, "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = LIBDIR / LIBNAME
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    # the proxy executable: host code over the C ABI, linked against the library
    exe = LIBDIR / PROXY
    exe_tmp = LIBDIR / (PROXY + ".tmp")
    cmd = [_hipcc(), "-O2", "-std=c++17", "-I", str(PKG.parent / "include"), str(HOST / "proxy.cpp"),
           "-L", str(LIBDIR), "-lseissol_amd", "-Wl,-rpath,$ORIGIN", "-o", str(exe_tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"proxy build failed\n{' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(exe_tmp, exe)
    if verbose:
        print(f"built {out} and {exe}")
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.j, verbose=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
