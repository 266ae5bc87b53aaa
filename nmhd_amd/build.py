"""Build libnmhd.so (HIP, gfx950) in-tree with explicit hipcc commands.

The library is the product: nmhd_amd/libnmhd.so, C ABI in include/nmhd.h.
Every HIP translation unit is compiled with packed-FP32 instruction
selection disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnmhd.so")

SOURCES = ["capi.hip", "capi_batch.hip", "capi_adjoint.hip", "kernels_common.hip",
           "kernels_unfused.hip", "kernels_strict.hip", "kernels_tiled.hip", "kernels_batch.hip",
           "kernels_adjoint.hip", "kernels_diag.hip", "capi_noalloc.hip"]
# kernels_strict.hip reproduces the reference's rounding: no FMA contraction.
FILE_FLAGS = {"kernels_strict.hip": ["-ffp-contract=off"]}

HIPFLAGS = [
    # -ffp-contract=on: FMA contraction only within one source expression, so
    # the same expression rounds the same way in every kernel that inlines it
    # (a cell's result does not depend on the tiling or the kernel variant)
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_cpp():
    """The C++ facade's programs: the example driver and the C++ tests."""
    inc = ["-I", os.path.join(ROOT, "include")]
    link = ["-L", HERE, "-lnmhd", "-L/opt/rocm/lib", "-Wl,-rpath-link,/opt/rocm/lib"]
    bindir = os.path.join(HERE, "bin")
    tbin = os.path.join(ROOT, "tests", "cpp", "bin")
    os.makedirs(bindir, exist_ok=True)
    os.makedirs(tbin, exist_ok=True)
    flags = ["g++", "-std=c++14", "-O2", "-Wall"]
    _run(flags + inc + [os.path.join(ROOT, "tests", "cpp", "crystal_test.cpp"),
                        "-o", os.path.join(tbin, "crystal_test")])
    _run(flags + inc + [os.path.join(ROOT, "tests", "cpp", "setup_format_test.cpp"),
                        "-o", os.path.join(tbin, "setup_format_test")])
    _run(flags + inc + [os.path.join(ROOT, "tests", "cpp", "facade_test.cpp"),
                        "-o", os.path.join(tbin, "facade_test")] + link +
         ["-Wl,-rpath,$ORIGIN/../../../nmhd_amd"])
    _run(flags + inc + [os.path.join(HERE, "apps", "kepler_disk.cpp"),
                        "-o", os.path.join(bindir, "kepler_disk")] + link +
         ["-Wl,-rpath,$ORIGIN/.."])
    # the VALU issue-rate probe that places the face kernel on its roofline
    _run(["hipcc", "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", os.path.join(HERE, "apps", "valu_rate.hip"),
          "-o", os.path.join(bindir, "valu_rate")])
    # host (double) build of the HLLD reverse sweep for the CPU derivative tests
    _run(["hipcc", "-x", "hip", "--offload-host-only", "-O2", "-std=c++17", "-fPIC", "-shared",
          "-I", CSRC, os.path.join(ROOT, "tests", "cpp", "adj_hlld_host.cpp"),
          "-o", os.path.join(tbin, "libadj_hlld_host.so")])


def build(jobs=8, extra=()):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append(["hipcc", *HIPFLAGS, *FILE_FLAGS.get(src, []), *extra, "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    _run(["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", LIB])
    build_cpp()
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
