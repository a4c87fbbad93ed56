"""Builds the native library (libtgp.so) and the `gpu_part` command in-tree.

hipcc compiles the gfx950 kernels and the HIP host driver; g++ compiles the
OpenMP workload generator.  Outputs land next to this file so they travel
with the source tree.  `python -m thegraphpartitioning_amd.build` rebuilds.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtgp.so")
CLI = os.path.join(PKG, "gpu_part")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include")]
# packed-fp32 instructions are kept out of every device code object
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = ["tgp_kernels.hip", "tgp_learn.hip", "tgp_balance.hip", "tgp_driver.cpp",
               "tgp_host.cpp"]
GXX_SOURCES = ["tgp_gen.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    return r.stdout


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs, steps = [], []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        steps.append([HIPCC] + DEVICE + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    for src in GXX_SOURCES:
        o = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
        steps.append(["g++"] + COMMON + ["-fopenmp", "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    main_o = os.path.join(OBJ, "gpu_part_main.o")
    steps.append([HIPCC] + COMMON + ["-x", "hip", "--offload-arch=" + ARCH, "-c",
                                     os.path.join(CSRC, "gpu_part_main.cpp"), "-o", main_o])
    # the compile steps are independent: run them side by side
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(len(steps), os.cpu_count() or 1, 8)) as ex:
        list(ex.map(lambda c: _run(c, verbose), steps))
    tmp = LIB + ".tmp"
    _run([HIPCC] + DEVICE + ["-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp", "-ldl"], verbose)
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run([HIPCC, "-o", tmp, main_o, "-L" + PKG, "-ltgp", "-Wl,-rpath,$ORIGIN", "-lgomp"], verbose)
    os.replace(tmp, CLI)
    return LIB


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    for p in (LIB, CLI):
        if os.path.exists(p):
            os.remove(p)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, "and", CLI)
