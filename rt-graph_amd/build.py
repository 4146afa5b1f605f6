"""Build recipe for librtg.so (HIP for gfx950) and the bs_tc command-line driver.

Every HIP translation unit is compiled by hipcc for gfx950 only, with packed FP32
instructions disabled, into _build/, then linked into librtg.so beside this file.
Run as a script or call build(); each call recompiles everything.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
HOST = HERE / "host"
BUILD = HERE / "_build"
LIB = HERE / "librtg.so"
CLI = HERE / "bs_tc"
ROOT = HERE.parent

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("RTG_CXX", "/opt/rocm/llvm/bin/clang++")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-Wall", "-Wno-unused-function", f"-I{ROOT / 'include'}"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc",
]

HIP_SOURCES = sorted(CSRC.glob("*.hip"))
CPP_SOURCES = sorted(CSRC.glob("*.cpp"))


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        # the host pass does not know the gfx target feature; harmless
        if "packed-fp32-ops" not in line and "not a recognized feature" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *COMMON, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [CXX, *COMMON, "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    if BUILD.exists():
        shutil.rmtree(BUILD)
    BUILD.mkdir(parents=True)
    sources = HIP_SOURCES + CPP_SOURCES
    jobs = jobs or min(8, os.cpu_count() or 4, len(sources))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs),
          "-lpthread"])
    os.replace(tmp, LIB)
    # Command-line driver mirroring tc/bs_tc/main.cu: bs_tc <edge_file> <device_id>
    cli_src = HOST / "bs_tc_main.cpp"
    if cli_src.exists():
        tmpc = CLI.with_suffix(".tmp")
        _run([CXX, *COMMON, f"-I{HOST}", str(cli_src), str(HOST / "graph.cpp"),
              "-o", str(tmpc), f"-L{HERE}", "-lrtg", f"-Wl,-rpath,{HERE}", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmpc, CLI)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
