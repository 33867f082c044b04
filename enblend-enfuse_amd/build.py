"""Build the gfx950 shared library of the enblend/enfuse pyramid path.

Compiles every HIP source under csrc/ with hipcc and the host-only C++
sources (the TIFF codec) with g++ (one object per source, in parallel, under
_build/) and links, with zlib,
enblend-enfuse_amd/libenblend_amd.so (in-tree, so it travels with the source
tree).  Flags:
  * --offload-arch=gfx950 only (MI355X);
  * -ffp-contract=off so that double arithmetic keeps the reference's
    operation order (no FMA contraction);
  * -Xclang -target-feature -Xclang -packed-fp32-ops keeps the compiler from
    forming packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libenblend_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def host_sources():
    """Host-only C++ (no device code): compiled with g++."""
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link the library;
    raises CalledProcessError on failure."""
    from concurrent.futures import ThreadPoolExecutor

    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(HERE, "..", "include")]
    compile_flags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        if src.endswith(".cpp"):
            cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall"] + inc + ["-c", src, "-o", obj]
        else:
            cmd = [HIPCC] + compile_flags + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(compile_one, sources() + host_sources()))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs + ["-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
