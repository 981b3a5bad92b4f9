"""Build libdmolp_hip.so for gfx950 in-tree (hipcc, no JIT cache).

Usage: python -m dmolp_amd.build [--jobs N] [--force]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libdmolp_hip.so"

SOURCES = [
    "graph.hip",
    "select.hip",
    "phase2.hip",
    "phase1.hip",
    "gen_dev.hip",
    "api.hip",
    "batch.hip",
    "server.hip",
    "halo.hip",
    "hubs.hip",
    "noalloc.hip",
    "opbatch.hip",
    "probe.hip",
    "gen_host.cpp",
    "textio.cpp",
    "driver/driver.cpp",
    "driver/report.cpp",
]
CLI = PKG / "hpc_partitioning"

ARCH = os.environ.get("DMOLP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # Parity arithmetic (penalties, ratio, scores) must be plain IEEE
    # operations in source order.
    "-ffp-contract=off",
]
HIP_ONLY = [
    "-x", "hip",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return path


def _obj(src: str) -> Path:
    return BUILD / (src.replace(".", "_").replace("/", "_") + ".o")


def _deps_hash(src: Path) -> str:
    h = hashlib.sha256()
    h.update(src.read_bytes())
    for hdr in sorted(CSRC.glob("*.hpp")) + sorted((ROOT / "include").rglob("*.h")):
        h.update(hdr.read_bytes())
    h.update(" ".join(COMMON + HIP_ONLY).encode())
    return h.hexdigest()


def compile_one(src: str, force: bool) -> Path:
    s = CSRC / src
    o = _obj(src)
    stamp = o.with_suffix(".hash")
    digest = _deps_hash(s)
    if not force and o.exists() and stamp.exists() and stamp.read_text() == digest:
        return o
    cmd = [hipcc()] + COMMON + (HIP_ONLY if src.endswith(".hip") else []) + [
        "-I", str(ROOT / "include"), "-c", str(s), "-o", str(o)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    stamp.write_text(digest)
    return o


def build(jobs: int | None = None, force: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, force), SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp)] + [
        str(o) for o in objs] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # The reference's command line over the drop-in run_phase1/run_phase2.
    cmd = [hipcc()] + COMMON + ["-I", str(ROOT / "include"), str(CSRC / "driver" / "main.cpp"),
                                "-o", str(CLI), "-L", str(PKG), "-ldmolp_hip",
                                "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return LIB


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.jobs, a.force))
    return 0


if __name__ == "__main__":
    sys.exit(main())
