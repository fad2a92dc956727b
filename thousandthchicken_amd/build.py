"""Builds libtc.so (HIP kernels for gfx950 + the C++ host path) in-tree.

Every translation unit is compiled by hipcc for gfx950 only; the shared library
lands next to this file so that it travels with the repository snapshot.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtc.so")
OBJ = os.path.join(HERE, "_objs")

SOURCES = [
    "device.hip",
    "kernels/t1_wave.hip",
    "kernels/t1_stripe.hip",
    "kernels/t1_lockstep.hip",
    "kernels/idwt.hip",
    "kernels/colour.hip",
    "kernels/fdwt.hip",
    "host/codestream.cpp",
    "host/pipeline.cpp",
    "host/plan_util.cpp",
    "host/decoder_api.cpp",
    "host/stages.cpp",
    "host/server.cpp",
    "host/forward.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I" + os.path.join(HERE, "..", "include"),
]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [HIPCC, "-x", "hip"] + FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


ENC_SRC = os.path.join(CSRC, "encoder", "encoder.cpp")
ENC_OUT = os.path.join(HERE, "libtc_encode.so")


def build_encoder():
    """Host-only encoder library (synthetic workloads for tests and benchmarks)."""
    tmp = ENC_OUT + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-o", tmp, ENC_SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"encoder build failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, ENC_OUT)
    return ENC_OUT


def build(verbose=False):
    """Compiles every source and links libtc.so (and the encoder).  Always rebuilds."""
    build_encoder()
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + [o for o, _ in results] + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    shutil.rmtree(OBJ, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
