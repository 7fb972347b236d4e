"""Builds map_amd/libmapx.so in-tree.

Host C++ (DAG, fuser, code generator, runtime, C ABI) is compiled with g++
against the HIP runtime headers; the ahead-of-time device code (input
generators) with hipcc for gfx950; the device template header used by the
run-time compiler (hiprtc) is embedded as a string. Run as
``python -m map_amd.build`` or through ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmapx.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HOST_SOURCES = ["dag.cpp", "fuser.cpp", "codegen.cpp", "runtime.cpp", "serve.cpp", "rasterio.cpp", "dev.cpp", "bindings.cpp"]
DEVICE_SOURCES = ["gen.hip", "radial.hip", "spread.hip", "mathcheck.hip"]
HIP_DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXXFLAGS = ["-O2", "-g0", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
            "-I" + os.path.join(ROCM, "include"), "-I" + BUILD, "-I" + CSRC]


def _embed_header():
    src = os.path.join(CSRC, "device", "mapx_dev.hiph")
    with open(src) as f:
        text = f.read()
    if ")MXHDR\"" in text:
        raise RuntimeError("device header contains the raw-string delimiter")
    out = os.path.join(BUILD, "mapx_dev_hiph.inc")
    body = "// generated from device/mapx_dev.hiph by map_amd/build.py\n" \
           "static const char kDevHeader[] = R\"MXHDR(" + text + ")MXHDR\";\n"
    old = open(out).read() if os.path.exists(out) else None
    if old != body:
        with open(out, "w") as f:
            f.write(body)


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    _embed_header()
    cmds, objs = [], []
    for s in HOST_SOURCES:
        o = os.path.join(BUILD, s.replace(".cpp", ".o"))
        cmds.append(["g++"] + CXXFLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    for s in DEVICE_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        cmds.append([os.path.join(ROCM, "bin", "hipcc"), "-O3", "-fPIC", "-std=c++17"] + HIP_DEVICE_FLAGS
                    + ["-c", os.path.join(CSRC, s), "-o", o])
        objs.append(o)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "-shared", "-fPIC", "-o", tmp] + objs
         + ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lhiprtc", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
