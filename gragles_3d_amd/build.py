"""Build libgg_levelset.so in-tree for gfx950.

Every HIP translation unit is compiled with hipcc for gfx950 only, with FMA
contraction off (the FP64 path must round exactly like the CPU restatement)
and packed-FP32 instruction selection disabled; the host set-up code is plain
C++.  Objects go to gragles_3d_amd/_build/, the shared library next to this
file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgg_levelset.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_SOURCES = ["gg_api.hip", "gg_plan.hip", "gg_conv.hip", "gg_compare.hip", "gg_comparison.hip", "gg_hull.hip",
               "gg_redist.hip", "gg_coarsen.hip", "gg_correct.hip"]
CXX_SOURCES = ["gg_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + INCLUDE,
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build(jobs=None, verbose=False, force=False):
    """Compile the HIP and C++ sources (only those newer than their objects
    unless force) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".h", ".inc"))]
    headers += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    hdr_mtime = max(os.path.getmtime(h) for h in headers)
    cmds = []
    objs = []
    for src in HIP_SOURCES + CXX_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(BUILD, src + ".o")
        objs.append(o)
        if not force and os.path.exists(o) and os.path.getmtime(o) >= max(os.path.getmtime(s), hdr_mtime):
            continue
        if src.endswith(".hip"):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o])
        else:
            cmds.append(["g++"] + CXX_FLAGS + ["-c", s, "-o", o])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


def build_abi_check(src, out):
    """Compile a C program against include/*.h and link it to the library (the
    run-time path of the library is recorded relative to the program)."""
    os.makedirs(os.path.dirname(out), exist_ok=True)
    rel = os.path.relpath(HERE, os.path.dirname(out))
    _run(["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I" + INCLUDE, src, "-o", out, "-L" + HERE, "-lgg_levelset",
          "-Wl,-rpath,$ORIGIN/" + rel])
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, force="-f" in sys.argv))
