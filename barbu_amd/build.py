"""Builds libftl.so (HIP kernels for gfx950 + the C ABI) in place.

    python -m barbu_amd.build           # or barbu_amd.build.build()

Every launch class is its own translation unit (ftl_plain.hip with
-DFTL_CLASS_ID=k: the plain kernels and the tile-list kernel of batched steps
of that class), so the objects compile in parallel.
"""
import concurrent.futures
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libftl.so")
NCLASSES = int(re.search(r"#define FTL_NCLASSES (\d+)", open(os.path.join(HERE, "csrc", "ftl_classes.h")).read()).group(1))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # Every fma in the kernels is written out; nothing else may be fused.
    "-ffp-contract=off",
    # No packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _units():
    units = []
    for k in range(NCLASSES):
        # The tile's loops are fully unrolled (#pragma unroll) so the strand
        # stays in registers; the 64-point class's bodies exceed LLVM's default
        # size limit for that, and would otherwise keep the points in scratch.
        # (The other classes compile to the same code with or without it.)
        units.append(("ftl_plain.hip", ["-DFTL_CLASS_ID=%d" % k, "-mllvm", "-pragma-unroll-threshold=200000"],
                      "plain_%d.o" % k))
    for src in ("ftl_misc.hip", "ftl_render.hip", "ftl_init.hip", "ftl_marschner.hip", "ftl_api.cpp", "ftl_batch_api.cpp", "ftl_autotune.cpp",
                "ftl_scenario.cpp", "ftl_server.cpp"):
        units.append((src, [], os.path.splitext(src)[0] + ".o"))
    return units


def _compile(unit):
    src, defs, obj = unit
    out = os.path.join(BUILD, obj)
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + COMMON + defs + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return out


def build(jobs=None, verbose=True):
    """Compiles every unit and links barbu_amd/libftl.so.  Always rebuilds."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(BUILD, exist_ok=True)
    units = _units()
    jobs = jobs or min(16, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built %s (%d units)" % (LIB, len(units)))
    return LIB


if __name__ == "__main__":
    build(jobs=int(sys.argv[1]) if len(sys.argv) > 1 else None)
