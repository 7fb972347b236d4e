"""Builds tests/host/_build/host_driver: the runtime's host sources (map_amd/csrc, all but the
HIP device layer dev.cpp) with the device double (dev_double.cpp) and the driver
(host_driver.cpp), under AddressSanitizer and UndefinedBehaviorSanitizer, with g++ only (no
device code). Run as ``python -m tests.host.build``; tests/test_host_sanitized.py builds and
runs it."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "map_amd", "csrc")
OUT = os.path.join(HERE, "_build")
EXE = os.path.join(OUT, "host_driver")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
HOST = ["dag.cpp", "fuser.cpp", "codegen.cpp", "runtime.cpp", "serve.cpp", "rasterio.cpp", "bindings.cpp"]


def build():
    sys.path.insert(0, ROOT)
    from map_amd import build as B
    os.makedirs(B.BUILD, exist_ok=True)
    B._embed_header()  # the device header text the double hands to the code generator's users
    os.makedirs(OUT, exist_ok=True)
    flags = ["g++", "-O1", "-g", "-std=c++17", "-Wall", "-Wno-unused-function"] + SAN + \
            ["-I" + B.BUILD, "-I" + CSRC]
    srcs = [os.path.join(CSRC, s) for s in HOST] + [os.path.join(HERE, "dev_double.cpp"),
                                                     os.path.join(HERE, "host_driver.cpp")]
    objs = [os.path.join(OUT, os.path.basename(s) + ".o") for s in srcs]

    def one(pair):
        src, obj = pair
        if os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), _newest_header()):
            return
        r = subprocess.run(flags + ["-c", src, "-o", obj], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError("host build failed: " + src + "\n" + r.stdout)

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(one, zip(srcs, objs)))
    r = subprocess.run(["g++"] + SAN + objs + ["-o", EXE, "-lpthread"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("host link failed\n" + r.stdout)
    return EXE


def _newest_header():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")]
    hs.append(os.path.join(ROOT, "include", "mapx.h"))
    return max(os.path.getmtime(h) for h in hs)


if __name__ == "__main__":
    print(build())
