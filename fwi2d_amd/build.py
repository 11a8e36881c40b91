"""Build the gfx950 shared library fwi2d_amd/libfwi2d_hip.so with hipcc.

    python -m fwi2d_amd.build          (or fwi2d_amd.build.build())

Every HIP source is compiled for gfx950 only, with packed-FP32 instructions
disabled, and linked in-tree so the library travels with the repository snapshot.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfwi2d_hip.so")

HIP_SOURCES = ["kernels.hip", "engine.hip", "acoustic.hip"]
CXX_SOURCES = ["host_op.cpp", "solver.cpp", "io.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(verbose=False):
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="fwi2d_build_")
    try:
        objs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            out = _run([cc, "-c", "-x", "hip"] + COMMON + HIP_FLAGS + [os.path.join(CSRC, src), "-o", obj])
            if verbose and out.strip():
                print(out)
            objs.append(obj)
        for src in CXX_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            _run(["g++", "-c"] + COMMON + ["-D__HIP_PLATFORM_AMD__", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        out_tmp = LIB + ".tmp"
        _run([cc, "-shared", "--offload-arch=gfx950", "-o", out_tmp] + objs)
        os.replace(out_tmp, LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
