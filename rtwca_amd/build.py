"""Build librtw.so (the gfx950 integrator behind include/rtw.h) in-tree.

    python -m rtwca_amd.build          # compile every .hip under csrc/ and link
    python -m rtwca_amd.build -D NAME=VALUE --out DIR/librtw.so   # a tuning variant

hipcc cross-compiles for gfx950 without a GPU.  Objects go to rtwca_amd/_build/,
the library to rtwca_amd/librtw.so.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librtw.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Packed FP32 instructions are disabled for all device code.
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def _compile(src, bdir=BUILD, extra=()):
    obj = os.path.join(bdir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + list(extra) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=False, jobs=None, defines=(), out=None):
    """Compile and link; `defines` are NAME=VALUE tuning macros, `out` the library path."""
    lib = out or LIB
    bdir = BUILD if out is None else os.path.join(os.path.dirname(os.path.abspath(lib)), "_build")
    extra = ["-D" + d for d in defines]
    os.makedirs(bdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda f: _compile(f, bdir, extra), srcs))
    tmp = lib + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="build librtw.so for gfx950")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="NAME=VALUE tuning macro")
    ap.add_argument("--out", default=None, help="library path (default rtwca_amd/librtw.so)")
    a = ap.parse_args()
    build(verbose=True, defines=a.defines, out=a.out)
    sys.exit(0)
