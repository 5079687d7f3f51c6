"""Build the gfx950 HIP library (and the CPU oracle) in-tree.

    python fyzzxyzzy-sdk_amd/build.py           # libfyzz_hip.so + oracle
    python fyzzxyzzy-sdk_amd/build.py --ref     # also the reference checker

The library is fyzzxyzzy-sdk_amd/fyzz_amd/libfyzz_hip.so. Device code is
compiled for gfx950 only, with floating-point contraction off (the
arithmetic order of the reference's FP32 paths is kept) and packed-FP32
instruction selection disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "fyzz_amd" / "libfyzz_hip.so"
OBJ = PKG / "_build"

HOST_SOURCES = [CSRC / "islands.cpp"]  # host-only parts of the library
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
DEVICE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++20", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources) or 1)) as ex:
        for src in sources:
            obj = OBJ / (src.stem + ".o")
            objs.append(obj)
            cmd = [HIPCC, *DEVICE_FLAGS, "-c", str(src), "-o", str(obj)]
            jobs.append(ex.submit(_run, cmd))
        for src in HOST_SOURCES:
            obj = OBJ / (src.stem + ".o")
            objs.append(obj)
            cmd = ["g++", "-O2", "-fPIC", "-std=c++20", f"-I{ROOT / 'include'}", "-c", str(src), "-o", str(obj)]
            jobs.append(ex.submit(_run, cmd))
        for j in jobs:
            j.result()
    scenes_so = PKG / "fyzz_amd" / "libfyzz_scenes.so"
    _run(["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-o", str(scenes_so), str(CSRC / "scenes_host.cpp")])
    tmp = OUT.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


def build_oracle(ref: bool = False) -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    if ref and Path("/root/reference/src/native/scene.cpp").exists():
        _run(["make", "-s", "-j8", "-C", str(ROOT / "oracle"), "ref"])


if __name__ == "__main__":
    build_hip(verbose=True)
    build_oracle(ref="--ref" in sys.argv)
