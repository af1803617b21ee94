"""Build recipes for the native libraries (in-tree, gfx950 only).

    python -m multians_amd.build          # product library + drop-in C++ layer + demo
    python -m multians_amd.build --tests  # also the CPU emulation used by tests

Outputs go to multians_amd/lib/.  hipcc cross-compiles for gfx950 without a GPU.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code (no float math here,
# but keep the compiler from forming v_pk_* from any host-style helper code).
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["mans_kernels.hip", "mans_tables.hip", "mans_encode.hip", "mans_interleaved.hip",
               "mans_runtime.hip",
               "mans_codec.cc", "mans_container.cc"]
CUHD_SOURCES = ["cuhd_gpu_decoder.cc", "cuhd_buffers.cc", "ans_compat.cc"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_product():
    os.makedirs(LIB, exist_ok=True)
    so = os.path.join(LIB, "libmultians_amd.so")
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-I" + INC]
         + [os.path.join(CSRC, s) for s in LIB_SOURCES] + ["-o", so])
    cuhd = [os.path.join(CSRC, s) for s in CUHD_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    if cuhd:
        # drop-in C++ classes: plain C++ over the C ABI (no HIP headers)
        _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + INC] + cuhd
             + ["-L" + LIB, "-lmultians_amd", "-Wl,-rpath,$ORIGIN", "-o",
                os.path.join(LIB, "libcuhd_amd.so")])
        demo = os.path.join(CSRC, "demo_main.cc")
        if os.path.exists(demo):
            _run(["g++", "-std=c++17", "-O2", "-pthread", "-I" + INC, demo, "-L" + LIB,
                  "-lcuhd_amd", "-lmultians_amd", "-Wl,-rpath,$ORIGIN", "-o",
                  os.path.join(LIB, "mans_demo")])
    return so


def build_test_support():
    os.makedirs(LIB, exist_ok=True)
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + INC,
          os.path.join(CSRC, "mans_emu.cc"), "-o", os.path.join(LIB, "libmans_emu.so")])


if __name__ == "__main__":
    build_product()
    if "--tests" in sys.argv:
        build_test_support()
    print("built", LIB)
