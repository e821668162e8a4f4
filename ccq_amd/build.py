"""Build the gfx950 engine library in-tree: ccq_amd/libccq_hip.so.

Steps: regenerate the embedded contraction IR from the spin-orbital programs
(ccq_amd/gen), then compile the HIP sources with hipcc for gfx950.
"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libccq_hip.so")
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("engine.hip", "kernels.hip", "guess.hip")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    # no packed FP32 (v_pk_*_f32) code generation anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def generate_ir(methods=None) -> None:
    sys.path.insert(0, ROOT)
    from ccq_amd.gen import emit
    emit.generate_all(os.path.join(HERE, "csrc"), methods or emit.METHODS)


def build(verbose: bool = False, methods=None) -> str:
    generate_ir(methods)
    cmd = [HIPCC] + FLAGS + SOURCES + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libccq_hip.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
