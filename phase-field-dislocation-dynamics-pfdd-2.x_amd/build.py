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


# Translation units of the HIP library, compiled in parallel and linked
# into one shared object (no device code crosses a unit boundary).
HIP_UNITS = ("pfdd_gl.hip", "pfdd_pass.hip", "pfdd_update.hip", "pfdd_mid.hip", "pfdd_plane.hip")


def build_lib():
    lib = os.path.join(HERE, "lib")
    obj = os.path.join(lib, "obj")
    os.makedirs(obj, exist_ok=True)
    out = os.path.join(lib, "libpfdd_gl.so")
    units = [u for u in HIP_UNITS if os.path.exists(os.path.join(HERE, "csrc", u))]
    cmds, objs = [], []
    # PFDD_INCREMENTAL=1 (development only) skips units whose object is newer
    # than the unit and every header; the default rebuilds everything.
    incremental = os.environ.get("PFDD_INCREMENTAL") == "1"
    headers = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc")) if f.endswith(".hpp")]
    headers.append(os.path.join(ROOT, "include", "pfdd_gl.h"))
    newest_hdr = max(os.path.getmtime(f) for f in headers)
    for u in units:
        o = os.path.join(obj, u.replace(".hip", ".o"))
        objs.append(o)
        src = os.path.join(HERE, "csrc", u)
        if incremental and os.path.exists(o) and os.path.getmtime(o) > max(newest_hdr, os.path.getmtime(src)):
            continue
        cmds.append([HIPCC] + HIPFLAGS + ["-c", "-o", o, src])
    procs = []
    for c in cmds:
        print(" ".join(c), flush=True)
        procs.append(subprocess.Popen(c))
    bad = [c for c, p in zip(cmds, procs) if p.wait() != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])
    run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
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
