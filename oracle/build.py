"""Builds the C/OpenMP CPU restatement (oracle/libtiled_cpu.so) with gcc.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline). The reference's
own sources are CUDA/C++ tied to nvcc and libtorch, so no reference code is
compiled here.
"""

from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "tiled_cpu.c"
LIB = HERE / "libtiled_cpu.so"
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-fno-math-errno", "-fno-trapping-math"]


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    proc = subprocess.run(["gcc", *FLAGS, str(SRC), "-o", str(tmp), "-lm"], capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{proc.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
