"""Build libmarvl_align.so (HIP for gfx950 + host C++) and the `daligner`
driver in place.  `python -m marvel_amd.build`"""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MI_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no packed FP32 instructions in any device code
DEV_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]

LIB_SOURCES = ["la_batch.hip", "la_batch_noalloc.hip", "la_gen.hip", "filter.hip", "trace_pts.hip", "scorer.hip", "scorer_train.hip", "dust.hip", "selftest.hip", "spec.cpp", "align_capi.cpp", "ovl_buffer.cpp",
               "db_block.cpp", "tracks.cpp"]
CLI_SOURCE = "daligner_main.cpp"


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in LIB_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + DEV_FLAGS + ["-I", os.path.join(HERE, "..", "include"), "-c", s, "-o", o]
        if src.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(HERE, "..", "include"),
                   "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(o)
    out = os.path.join(HERE, "libmarvl_align.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, out)
    # the GPU daligner driver, linked against the library next to it
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    exe = os.path.join(bindir, "daligner")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"),
          os.path.join(CSRC, CLI_SOURCE), "-o", exe + ".tmp", "-L", HERE, "-lmarvl_align",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(exe + ".tmp", exe)
    # the .las utilities: LAsort and LAmerge over the library's GPU sort, LAcat (host only)
    for name, src, lib in (("LAsort", "lasort_main.cpp", True), ("LAmerge", "lamerge_main.cpp", True),
                           ("DBdust", "dbdust_main.cpp", True),
                           ("LAcat", "lacat_main.cpp", False)):
        exe = os.path.join(bindir, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(HERE, "..", "include"), os.path.join(CSRC, src),
               "-o", exe + ".tmp"]
        if lib:
            cmd += ["-L", HERE, "-lmarvl_align", "-Wl,-rpath,$ORIGIN/.."]
        _run(cmd)
        os.replace(exe + ".tmp", exe)
    # the block-pair plan emitter (host only)
    hpc = os.path.join(bindir, "HPCdaligner")
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "hpcdaligner_main.cpp"), "-o", hpc + ".tmp"])
    os.replace(hpc + ".tmp", hpc)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
