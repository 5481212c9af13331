"""Build the two native libraries of the MI355X backend in-tree.

* ``libgrico_host.so`` -- host plan compiler (C++17, g++).
* ``libgrico_mf.so``   -- HIP kernels for gfx950 plus the device C ABI (hipcc).

Both land next to this file in ``lib/``. ``build()`` always recompiles; the
loaders never build anything themselves.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIBDIR = os.path.join(HERE, "lib")
HOST_SRC = [os.path.join(HERE, "csrc", "host", f) for f in ("topology.cpp", "space.cpp", "transfer.cpp", "capi.cpp")]
DEVICE_DIR = os.path.join(HERE, "csrc", "device")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_host():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libgrico_host.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-I", os.path.join(ROOT, "include"), "-o", out] + HOST_SRC)
    return out


def device_sources():
    return sorted(os.path.join(DEVICE_DIR, f) for f in os.listdir(DEVICE_DIR) if f.endswith(".hip"))


def build_device():
    """One object per source, compiled in parallel, then one shared-library link."""
    from concurrent.futures import ThreadPoolExecutor

    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    out = os.path.join(LIBDIR, "libgrico_mf.so")
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    jobs = []
    for src in device_sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        jobs.append((src, obj))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(workers) as pool:
        list(pool.map(lambda j: _run([HIPCC] + compile_flags + ["-c", "-o", j[1], j[0]]), jobs))
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + [o for _, o in jobs] +
         ["-lrocsolver", "-lrocblas"])
    return out


def build():
    build_host()
    build_device()


if __name__ == "__main__":
    what = sys.argv[1:] or ["host", "device"]
    if "host" in what:
        build_host()
    if "device" in what:
        build_device()
