"""Build the gfx950 engine library in-tree: ccq_amd/libccq_hip.so.

Steps: regenerate the embedded contraction IR from the spin-orbital programs
(ccq_amd/gen), compile each HIP source to an object with hipcc for gfx950
(in parallel), then link the shared library.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libccq_hip.so")
OBJDIR = os.path.join(HERE, "_build")
# the GEMM kernel instantiations are split by arithmetic / storage pair
# (gemm_f64, gemm_f32, gemm_f32io) so that they compile in parallel
SOURCES = [os.path.join(HERE, "csrc", f) for f in ("engine.hip", "kernels.hip", "guess.hip", "gemm_f64.hip",
                                                    "gemm_f32.hip", "gemm_f32io.hip")]
# host-only sources (no device code): built with the host compiler
HOST_SOURCES = [os.path.join(HERE, "csrc", f) for f in ("ints_io.cpp", "restart.cpp")]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 (v_pk_*_f32) code generation anywhere in this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def generate_ir(methods=None) -> None:
    sys.path.insert(0, ROOT)
    from ccq_amd.gen import emit
    emit.generate_all(os.path.join(HERE, "csrc"), methods or emit.METHODS)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd[-3:]))


def build(verbose: bool = False, methods=None) -> str:
    generate_ir(methods)
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.basename(s).replace(".hip", ".o")) for s in SOURCES]
    hobjs = [os.path.join(OBJDIR, os.path.basename(s).replace(".cpp", ".o")) for s in HOST_SOURCES]
    with cf.ThreadPoolExecutor(len(SOURCES) + len(HOST_SOURCES)) as ex:
        futs = [ex.submit(_run, [HIPCC] + FLAGS + ["-c", s, "-o", o], verbose)
                for s, o in zip(SOURCES, objs)]
        futs += [ex.submit(_run, [CXX, "-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-c", s,
                                  "-o", o], verbose) for s, o in zip(HOST_SOURCES, hobjs)]
        for f in futs:
            f.result()
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread"] + objs + hobjs +
         ["-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
