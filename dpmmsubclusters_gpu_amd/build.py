"""Builds the HIP library lib/libdpmm_hip.so for gfx950 (in-tree, so it travels with the repo).

    python -m dpmmsubclusters_gpu_amd.build        # or build() from Python
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libdpmm_hip.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for device code: some of their forms have given wrong
# results on MI355X boxes, silently.
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
                "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]

DEVICE_SOURCES = ["dpmm_device.hip", "inst_gauss_small.hip", "inst_gauss_64.hip", "inst_gauss_96.hip",
                  "inst_gauss_128.hip", "inst_mixed.hip", "dpmm_screener.hip"]
HOST_SOURCES = ["sampler.cpp"]
# The BF16 screens read every MFMA result back into VALU arithmetic: the VGPR form of the MFMA
# instructions saves one v_accvgpr_read per result element.
# dpmm_device.hip: the register forms of the FP64 triangular factorisations (D <= 128) are fully
# unrolled loops over register arrays; above LLVM's default unroll budget the arrays stay in memory
SOURCE_FLAGS = {"inst_mixed.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
                "dpmm_device.hip": ["-mllvm", "-pragma-unroll-threshold=200000"]}


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    objs = []
    procs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmd = [HIPCC] + DEVICE_FLAGS + SOURCE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        cmd = ["g++"] + HOST_FLAGS + ["-I", os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    for cmd, p in procs:
        if verbose:
            print(" ".join(cmd), flush=True)
        if p.wait() != 0:
            raise RuntimeError("compile failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp] + objs + ["-fopenmp", f"-L{ROCM}/lib", "-lamdhip64",
                                                 f"-Wl,-rpath,{ROCM}/lib"])
    shutil.move(tmp, LIB)
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    link = [f"-L{LIB_DIR}", "-ldpmm_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"]
    # the command-line front end (the reference's main: --params_path, --model_path, --result_path)
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "dpmm_main.cpp"), "-o",
          os.path.join(LIB_DIR, "dpmm_main")] + link)
    # the C++ drop-in header's example program (run by the GPU tests), linked against the library
    src = os.path.join(repo, "tests", "cpp", "dropin_readme.cpp")
    if os.path.exists(src):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(repo, "include"), src, "-o",
              os.path.join(LIB_DIR, "dropin_readme")] + link)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
