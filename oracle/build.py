"""Compile the CPU oracle (test infrastructure) into oracle/libldpc_oracle.so.

The reference itself (CUDA C++: nvcc, the CUDA runtime and cuRAND) cannot be
compiled in this image, so there is no oracle/_ref build; the oracle is the C
restatement in ldpc_oracle.c.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "libldpc_oracle.so"


CC = os.environ.get("CC", "gcc")
CFLAGS = ["-O3", "-std=c11", "-fPIC", "-shared", "-pthread", "-Wall", "-Wextra"]


def compiler_version() -> str:
    """First line of the C compiler's --version (recorded beside CPU timings)."""
    try:
        r = subprocess.run([CC, "--version"], capture_output=True, text=True)
        return r.stdout.splitlines()[0].strip() if r.returncode == 0 and r.stdout else CC
    except OSError:
        return CC


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [CC, *CFLAGS, str(HERE / "ldpc_oracle.c"), "-o", str(tmp), "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
