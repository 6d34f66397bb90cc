"""Build recipe for libgomc_hip.so (gfx950) and the CPU oracle library.

    python -m gomc_amd.build            # product library
    python -m gomc_amd.build --oracle   # also oracle/liboracle.so (tests)

Every HIP translation unit is compiled with hipcc for gfx950 only, without
FMA contraction (the FP64 paths keep the reference's operation order) and
with packed FP32 instructions disabled.
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
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libgomc_hip.so"
OBJ = PKG / "_obj"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]
SOURCES = ["box.hip", "box_setup.cpp", "driver.cpp", "mp.hip", "gcmc.cpp", "batch.hip", "batch_setup.cpp", "server.cpp",
           "placement.cpp", "gcmc_dev.hip", "virial.hip",
           "volume.cpp", "checkpoint.cpp"]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    if r.stderr.strip():
        lines = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build_library(jobs: int | None = None) -> Path:
    """Compile every source and link libgomc_hip.so in the package dir."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    OBJ.mkdir(exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def build_oracle() -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if Path("/root/reference").is_dir():
        r = subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference KAT build failed:\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "liboracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    args = ap.parse_args()
    print(build_library(args.j))
    if args.oracle:
        print(build_oracle())


if __name__ == "__main__":
    main()
