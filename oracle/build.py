"""Build recipe for the oracle (test infrastructure only).

* ``liboracle.so``  -- the C restatement of the reference step
  (nmhd_oracle.c), built with gcc, OpenMP and ``-ffp-contract=off``.
* ``_ref/libref_hlld.so`` -- the reference's own ``library/hlld.inl`` compiled
  for the host from the reference tree (``$NMHD_REF`` or /root/reference),
  used to pin the oracle's HLLD bit for bit.  Skipped when the tree is absent
  (a machine without it uses the built file, which travels with the
  repository's working tree).

Run ``python oracle/build.py`` or call :func:`build`.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DEFAULT = "/root/reference"


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_oracle():
    out = os.path.join(HERE, "liboracle.so")
    _run(["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
          "-std=c99", "-Wall", "-Wno-unknown-pragmas",
          os.path.join(HERE, "nmhd_oracle.c"), "-o", out, "-lm"])
    return out


def build_ref(ref_root=None):
    ref_root = ref_root or os.environ.get("NMHD_REF", REF_DEFAULT)
    src = os.path.join(ref_root, "src")
    if not os.path.isfile(os.path.join(src, "library", "hlld.inl")):
        print("reference tree not found at %s; oracle/_ref not rebuilt" % ref_root)
        return None
    outdir = os.path.join(HERE, "_ref")
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libref_hlld.so")
    _run(["hipcc", "-x", "hip", "--offload-host-only", "-O2",
          "-ffp-contract=off", "-fPIC", "-shared", "-I", src,
          os.path.join(HERE, "ref_hlld_driver.cpp"), "-o", out])
    return out


def build():
    build_oracle()
    build_ref()


if __name__ == "__main__":
    build()
    sys.exit(0)
