"""Build the gfx950 shared library and the sweep driver.

    python -m fast_ldpc_decoder_for_gpu_fixed_amd.build

The product is ``libldpc_mi355x.so`` next to this file: HIP kernels for gfx950
plus the C-ABI of ``include/ldpc_mi355x.h``.  Objects are compiled in
parallel with ``hipcc`` and linked in-tree so the library travels with the
source tree.  ``ldpc_sim`` beside it is the Eb/N0 sweep driver (the
reference's ``src/main.cpp``), a host program on that C-ABI.
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
BUILD = ROOT / "_build"
LIB = PKG / "libldpc_mi355x.so"
SIM = PKG / "ldpc_sim"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -Xclang -target-feature -Xclang -packed-fp32-ops keeps hipcc from forming
# packed fp32 instructions (the host pass warns that it does not know the
# feature; harmless).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = ["decode_kernels.hip", "sim_kernels.hip", "learn_kernels.hip", "pool_kernels.hip", "api.hip",
           "learn_api.hip", "pool_api.hip", "mixed_api.hip", "code.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        msg = "\n".join(line for line in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in line)
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{msg}")


def build_lib() -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    jobs = []
    for src in SOURCES:
        obj = BUILD / (src.replace(".", "_") + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        jobs.append(["hipcc" if HIPCC == "hipcc" else HIPCC, *HIP_FLAGS, *lang, "-c",
                     str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(_run, jobs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, LIB)
    build_sim()
    return LIB


def build_sim() -> Path:
    """The sweep driver: plain C++ on the C-ABI, linked to the in-tree library."""
    tmp = SIM.with_suffix(".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(CSRC / "ldpc_sim.cpp"),
          f"-L{PKG}", "-lldpc_mi355x", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, SIM)
    return SIM


def main(argv: list[str]) -> None:
    print("built", build_lib())


if __name__ == "__main__":
    main(sys.argv)
