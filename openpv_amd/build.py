"""Builds libpvhip.so (HIP, gfx950), libpvhost.so (C++ host) and the pvrun
executable in-tree.

    python -m openpv_amd.build          # the libraries and pvrun
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")
HOSTSRC = os.path.join(HERE, "host")

HIP_SOURCES = ["device.hip", "layer.hip", "recv_post.hip", "recv_pre.hip", "recv_pre_dense.hip", "hebbian.hip",
               "adjoint.hip"]
HOST_SOURCES = ["geometry.cpp", "params.cpp"]

# -ffp-contract=off keeps mul and add separately rounded in the EXACT kernels
# (FMA and MFMA paths call the builtins explicitly).  Packed FP32 instructions
# are disabled for all device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found; libpvhip cannot be built")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), res.stdout))


def build_hip(verbose=False):
    """Each source compiled to an object in parallel (device code is not
    relocatable: every object carries its own code object), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    hipcc = _hipcc()
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    objs = [os.path.join(objdir, os.path.splitext(src)[0] + ".o") for src in HIP_SOURCES]
    cmds = [[hipcc] + compile_flags + ["-I" + INCLUDE, "-c", "-o", o, os.path.join(CSRC, src)]
            for src, o in zip(HIP_SOURCES, objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda c: _run(c, verbose), cmds))
    out = os.path.join(LIBDIR, "libpvhip.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    return out


def build_host(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libpvhost.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + INCLUDE,
           "-o", tmp] + [os.path.join(HOSTSRC, s) for s in HOST_SOURCES]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build_pvrun(verbose=False):
    """The pvrun executable (runs a .params file) next to the libraries it links."""
    out = os.path.join(LIBDIR, "pvrun")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, "-o", tmp,
           os.path.join(HOSTSRC, "pvrun.cpp"), "-L" + LIBDIR, "-lpvhip", "-lpvhost",
           "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose=False):
    hip, host = build_hip(verbose), build_host(verbose)
    return hip, host, build_pvrun(verbose)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
