"""Builds libchili_gpu.so (gfx950 HIP kernels + host C++ + the C-ABI) in-tree.

    python -m chili_amd.build            # compile every source, link the library
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchili_gpu.so")
ARCH = "gfx950"

SOURCES = ["mc.hip", "vegas.hip", "flow.hip", "channels.cpp", "adapt.cpp", "common.cpp", "jit.cpp", "checkpoint.cpp", "optimize.cpp"]

# Device headers compiled again at run time (jit.cpp): embedded as text.
JIT_HEADERS = [("mc_device.hpp", "mc_device.hpp"), ("mc_kernels.hpp", "mc_kernels.hpp"), ("phys.hpp", "phys.hpp"),
               ("chanprog.hpp", "chanprog.hpp"), ("../../include/chili_gpu.h", os.path.join("..", "..", "include", "chili_gpu.h"))]


def _write_jit_headers():
    os.makedirs(BUILD, exist_ok=True)
    out = ["// generated by chili_amd/build.py: device headers for run-time compilation",
           "struct ChiliJitHeader { const char* name; const char* text; };",
           "static const ChiliJitHeader kJitHeaders[] = {"]
    for name, rel in JIT_HEADERS:
        with open(os.path.join(CSRC, rel)) as fh:
            text = fh.read()
        if ")CHILIJIT" in text:
            raise RuntimeError(f"{rel}: contains the raw-string delimiter")
        out.append(f'    {{"{name}", R"CHILIJIT({text})CHILIJIT"}},')
    out.append("};")
    out.append(f"static const int kJitHeaderCount = {len(JIT_HEADERS)};")
    path = os.path.join(BUILD, "jit_headers.inc")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        with open(path, "w") as fh:
            fh.write(new)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-lambda-capture"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if src.endswith(".hip") else ["-x", "c++"]) + \
        ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        # host-only translation units still need the HIP headers
        cmd = [HIPCC] + COMMON + ["-I", BUILD, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    _write_jit_headers()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
