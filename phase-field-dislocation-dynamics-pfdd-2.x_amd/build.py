"""Compile the gfx950 HIP library lib/libpfdd_gl.so (and the C++ host layer).

hipcc is driven directly (no torch JIT cache), output in-tree so the built
library ships with the source tree.  All HIP code is built with packed FP32
instructions disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("PFDD_OFFLOAD_ARCH", "gfx950")
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-Wno-unused-result"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    os.makedirs(os.path.join(HERE, "lib"), exist_ok=True)
    out = os.path.join(HERE, "lib", "libpfdd_gl.so")
    srcs = [os.path.join(HERE, "csrc", "pfdd_gl.hip")]
    run([HIPCC] + HIPFLAGS + ["-shared", "-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_host():
    out = os.path.join(HERE, "lib", "libpfdd_host.so")
    srcs = [os.path.join(HERE, "host", f) for f in ("pfdd_host.cpp", "pfdd_host_capi.cpp")]
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-o", out + ".tmp"] + srcs)
    os.replace(out + ".tmp", out)
    return out


def build_driver():
    """pfdd_hip: the deck driver (host C++) linked against the HIP library."""
    out = os.path.join(HERE, "lib", "pfdd_hip")
    srcs = [os.path.join(HERE, "host", f) for f in ("main.cpp", "pfdd_driver.cpp", "pfdd_host.cpp")]
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    run(["g++", "-O2", "-std=c++17", "-Wall", "-o", out + ".tmp"] + srcs +
        ["-L" + os.path.join(HERE, "lib"), "-lpfdd_gl", "-Wl,-rpath,$ORIGIN",
         "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(rocm, "lib")])
    os.replace(out + ".tmp", out)
    return out


def main():
    build_lib()
    build_host()
    build_driver()


if __name__ == "__main__":
    sys.exit(main())
