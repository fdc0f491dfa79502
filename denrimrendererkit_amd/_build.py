"""Builds the gfx950 shared library libdenrim_hip.so in-tree with hipcc.

Each source is compiled to an object in parallel, then linked. A second
library, libdenrim_hip_lds4.so, is the same code with 4 LDS traversal-stack
entries per lane instead of 24, so that tests drive the per-lane global
spill of the traversal stack on small scenes; the product loads the first.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libdenrim_hip.so")
LIB_LDS4 = os.path.join(PKG, "libdenrim_hip_lds4.so")
OBJ = os.path.join(os.path.dirname(PKG), "_build")
SOURCES = ["dr_api.cpp", "dr_bvh.cpp", "dr_sdf_build.cpp", "dr_server.cpp", "dr_unet.cpp", "dr_kernels.hip",
           "dr_denoise.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # IEEE evaluation with no contraction: camera rays and the triangle test
    # must produce the same bits as the reference formulas.
    "-ffp-contract=off",
    # No packed-FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _compile(src, obj, defines):
    cmd = [HIPCC, *FLAGS, *defines, "-c", "-o", obj, os.path.join(CSRC, src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed compiling {src}")
    return obj


def _library(out, tag, defines, pool):
    d = os.path.join(OBJ, tag)
    os.makedirs(d, exist_ok=True)
    objs = [os.path.join(d, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    return [pool.submit(_compile, s, o, defines) for s, o in zip(SOURCES, objs)], objs, out


def build(verbose: bool = False, variants: bool = True) -> str:
    jobs = [("default", LIB, [])]
    if variants:
        jobs.append(("lds4", LIB_LDS4, ["-DDR_LDS_STACK=4"]))
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        pending = [_library(out, tag, defs, pool) for tag, out, defs in jobs]
        for futures, objs, out in pending:
            for f in futures:
                f.result()
            cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp", *objs, "-lpthread"]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed linking {os.path.basename(out)}")
            os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
