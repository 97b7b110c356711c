"""Builds libastro_hot.so (the HIP hot path, gfx950) in-tree.

    python astrophage-microscope_amd/build.py [--jobs N]

Each .hip translation unit is compiled by hipcc to an object in parallel (the
host-only scenario layer by g++), then linked into
astrophage-microscope_amd/lib/libastro_hot.so. Flags:
  * --offload-arch=gfx950: MI355X only;
  * -ffp-contract=off: expressions round as written, so results compare with
    the CPU restatement of the reference bit for bit wherever no libm call
    intervenes;
  * -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 math.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libastro_hot.so"
SOURCES = ["world.hip", "kernels_hash.hip", "kernels_cells.hip", "kernels_fields.hip", "kernels_tau.hip", "surrogate.hip",
           "hash_noalloc.hip"]
# Host-only C++ above the C ABI (no device code), compiled by g++.
HOST_SOURCES = ["scenario.cpp"]
HOST_CXX = os.environ.get("CXX_HOST", "g++")
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall", "-Wextra",
              "-I", str(ROOT / "include"), "-I", str(CSRC)]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", str(ROOT / "include"),
    "-I", str(CSRC),
]


def _compile(src: str) -> tuple[str, int, str]:
    obj = BUILD / (Path(src).stem + ".o")
    if src in HOST_SOURCES:
        cmd = [HOST_CXX, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *FLAGS, "-c", str(CSRC / src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stdout + p.stderr


def build(jobs: int = 4, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        results = list(ex.map(_compile, SOURCES + HOST_SOURCES))
    failed = False
    for src, rc, out in results:
        noisy = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if rc != 0 or (verbose and noisy):
            print(f"--- {src} (rc={rc})\n" + "\n".join(noisy), file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("hipcc failed; see messages above")
    objs = [str(BUILD / (Path(s).stem + ".o")) for s in SOURCES + HOST_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    tmp.replace(LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
