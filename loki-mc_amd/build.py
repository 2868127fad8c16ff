"""Build step of the engine: compiles the gfx950 kernels and the C++ host
engine into loki-mc_amd/lib/libekmc.so with hipcc, and the command-line
driver (setup and LXCat readers, output writers) into
loki-mc_amd/bin/lokimc_gpu with g++ against that library.

    python loki-mc_amd/build.py

Packed-FP32 instructions are disabled for all device code
(-Xclang -target-feature -Xclang -packed-fp32-ops); the FP32 arithmetic of
the mixed form is issued as ordinary VALU operations.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
HOST_DIR = os.path.join(HERE, "host")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
HOST_SOURCES = ["setup_file.cpp", "gas_model.cpp", "outputs.cpp", "lokimc_gpu.cpp"]
OBJ_DIR = os.path.join(LIB_DIR, "obj")
ARCH = os.environ.get("EKMC_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
SOURCES = ["kernels.hip", "tables.cpp", "engine.cpp"]
# FP32 division and square root of the mixed form without IEEE rounding
# (a few ulp); FP64 arithmetic is unaffected.
EXTRA = {"kernels.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]}


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    for src in SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc] + HIP_FLAGS + EXTRA.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd = [hipcc, "-x", "hip"] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        objs.append(obj)
    out = os.path.join(LIB_DIR, "libekmc.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    build_cli(verbose)
    return out


def build_cli(verbose=False):
    """The command-line driver: host C++ only, linked against libekmc.so."""
    os.makedirs(BIN_DIR, exist_ok=True)
    exe = os.path.join(BIN_DIR, "lokimc_gpu")
    tmp = exe + ".tmp"
    cmd = (["g++", "-std=c++17", "-O2", "-Wall", "-I", INCLUDE] +
           [os.path.join(HOST_DIR, s) for s in HOST_SOURCES] +
           ["-L", LIB_DIR, "-lekmc", "-Wl,-rpath,$ORIGIN/../lib", "-L/opt/rocm/lib",
            "-Wl,-rpath-link,/opt/rocm/lib", "-o", tmp])
    _run(cmd, verbose)
    os.replace(tmp, exe)
    return exe


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    text = "\n".join(l for l in r.stdout.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), text))
    if verbose and text.strip():
        print(text)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
