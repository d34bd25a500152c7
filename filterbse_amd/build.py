"""Build the gfx950 library (libfbse_gpu.so) and the host-side C library
(libfbse_host.so) in-tree with hipcc / gcc.

    python -m filterbse_amd.build          # or filterbse_amd.build.build()
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

GPU_LIB = os.path.join(HERE, "libfbse_gpu.so")
HOST_LIB = os.path.join(HERE, "libfbse_host.so")

HIP_SOURCES = ["fft_kernels.hip", "fft_tile_kernels.hip", "ham_kernels.hip", "capi.hip", "mlp_kernels.hip",
               "build_kernels.hip", "io.hip", "gemm_kernels.hip"]
HOST_SOURCES = ["fbse_host.c"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 (v_pk_*_f32) is disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def build_gpu(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    objs = []
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        for src in HIP_SOURCES:
            obj = os.path.join(OBJ, src.replace(".hip", ".o"))
            objs.append(obj)
            cmd = [hipcc] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
            jobs.append(ex.submit(_run, cmd, verbose))
        for j in jobs:
            j.result()
    tmp = GPU_LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    _run(cmd, verbose)
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host(verbose=False):
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    tmp = HOST_LIB + ".tmp"
    cmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
           "-I" + os.path.join(ROOT, "include"), "-o", tmp] + srcs + [
           "-L" + HERE, "-lfbse_gpu", "-Wl,-rpath,$ORIGIN", "-lm", "-lgomp"]
    _run(cmd, verbose)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def build(verbose=False):
    build_gpu(verbose)
    build_host(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", GPU_LIB)
