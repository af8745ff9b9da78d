"""Builds the native libraries in-tree: libkyto_gpu.so (HIP for gfx950 + the host
C++ around it) and libkyto_synth.so (host C, the synthetic-data generator).

    python -m kyto_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
GPU_LIB = os.path.join(HERE, "libkyto_gpu.so")
SYNTH_LIB = os.path.join(HERE, "libkyto_synth.so")
# The sanitizer build (python -m kyto_amd.build sanitize): the host code of every
# source with AddressSanitizer and UndefinedBehaviorSanitizer (the device code as
# always), for CPU test runs (tests/sanitize.sh).
SAN_OBJ = os.path.join(HERE, "_build_san")
SAN_GPU_LIB = os.path.join(HERE, "libkyto_gpu_san.so")
SAN_SYNTH_LIB = os.path.join(HERE, "libkyto_synth_san.so")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
            "-Xarch_host", "-shared-libsan", "-g"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANG = os.path.join(ROCM, "lib", "llvm", "bin", "clang")

# Packed FP32 instructions are disabled (-packed-fp32-ops); no contraction, so every
# f64 sum the exact keys make is evaluated as written (fused multiply-adds are
# written out where wanted).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
# MFMA kernels keep their accumulators in ordinary VGPRs (no accumulator-register
# copies before the epilogue reads them).
# knn_bf16.hip compares keys that are never NaN: no NaN quieting before v_min_f32.
# Host code (the dense FP64 steps) for AVX2 hosts; no contraction there either.
HOST_FLAGS = {"dense.cpp": ["-Xarch_host", "-march=x86-64-v3"]}
MFMA_FLAGS = {
    "knn_bf16.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1", "-fno-honor-nans", "-mno-amdgpu-ieee"],
    "router.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
}
GPU_SOURCES = ["pca.hip", "knn.hip", "knn_bf16.hip", "router.hip", "ivf.hip", "sort.hip", "capi.cpp", "dense.cpp", "memory.cpp", "layer.hip", "connectivities.hip", "store.cpp", "dense_dev.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _compile(src: str, obj: str = OBJ, extra=()) -> str:
    out = os.path.join(obj, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    # .cpp sources hold host code only; hipcc compiles them with the HIP headers.
    cmd = [HIPCC, *HIP_FLAGS, *MFMA_FLAGS.get(src, []), *HOST_FLAGS.get(src, []), *extra, "-c", path, "-o", out]
    _run(cmd)
    return out


def build_gpu() -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(GPU_SOURCES))) as pool:
        objs = list(pool.map(_compile, GPU_SOURCES))
    tmp = GPU_LIB + ".tmp"
    # zstd (store chunks) is the system's libzstd.so.1, declared in store.cpp.
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread", "-l:libzstd.so.1"])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_synth() -> str:
    cc = shutil.which("gcc") or "cc"
    tmp = SYNTH_LIB + ".tmp"
    _run([cc, "-O3", "-fPIC", "-fopenmp", "-march=x86-64-v3", "-std=c11", "-shared",
          os.path.join(CSRC, "synth.c"), "-o", tmp, "-lm"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_sanitized() -> None:
    os.makedirs(SAN_OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(GPU_SOURCES))) as pool:
        objs = list(pool.map(lambda src: _compile(src, SAN_OBJ, SAN_HOST), GPU_SOURCES))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address",
          "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-shared-libsan", "-o", SAN_GPU_LIB, *objs,
          "-lpthread", "-l:libzstd.so.1"])
    _run([CLANG, "-O1", "-g", "-fPIC", "-march=x86-64-v3", "-std=c11", "-shared", "-Wno-unknown-pragmas",
          "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-shared-libsan",
          os.path.join(CSRC, "synth.c"), "-o", SAN_SYNTH_LIB, "-lm"])


def build_all() -> None:
    build_synth()
    build_gpu()


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "sanitize":
        build_sanitized()
        print("built", SAN_GPU_LIB, SAN_SYNTH_LIB, file=sys.stderr)
    else:
        build_all()
        print("built", GPU_LIB, SYNTH_LIB, file=sys.stderr)
