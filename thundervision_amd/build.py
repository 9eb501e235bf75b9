"""Builds the native libraries of thundervision_amd in-tree.

libtdvh.so     HIP kernels + C ABI (include/tdvh.h), gfx950, hipcc
libtdvsynth.so seeded synthetic stereo-pair generator (plain C, gcc)
libtdvision_mi355x.so C++ stage layer (include/tdvision/*.hpp) over libtdvh.so
_bin/test_stages  C++ tests of the stage layer (run by tests/test_stages_cpp.py)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["ctx.hip", "cost_bt.hip", "cost_window.hip", "dp.hip", "sgm.hip", "sgm_ref.hip", "sgm_wide.hip",
               "wta.hip", "batch.hip", "io.hip", "pipeline.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU forms stay disabled (none of these kernels need them)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + HIP_FLAGS)
    jobs = min(len(cmds), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(PKG, "libtdvh.so")
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", lib + ".tmp"] + objs)
    os.replace(lib + ".tmp", lib)
    # C++ stage layer: host-only code over the C ABI
    stage_srcs = [os.path.join(CSRC, "stages", f)
                  for f in ("runtime.cpp", "images.cpp", "stages.cpp", "matchers.cpp")]
    cxx = ["g++", "-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
           "-I" + INCLUDE, "-I/opt/rocm/include"]
    stlib = os.path.join(PKG, "libtdvision_mi355x.so")
    _run(cxx + ["-shared", "-o", stlib + ".tmp"] + stage_srcs +
         ["-L" + PKG, "-ltdvh", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath,/opt/rocm/lib", "-pthread"])
    os.replace(stlib + ".tmp", stlib)
    os.makedirs(os.path.join(PKG, "_bin"), exist_ok=True)
    tbin = os.path.join(PKG, "_bin", "test_stages")
    _run(cxx + ["-o", tbin + ".tmp", os.path.join(ROOT, "tests", "cpp", "test_stages.cpp"),
                "-L" + PKG, "-ltdvision_mi355x", "-ltdvh", "-L/opt/rocm/lib", "-lamdhip64",
                "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-pthread"])
    os.replace(tbin + ".tmp", tbin)
    synth = os.path.join(PKG, "libtdvsynth.so")
    _run(["gcc", "-O2", "-shared", "-fPIC", "-o", synth + ".tmp",
          os.path.join(CSRC, "synth.c"), "-lm"])
    os.replace(synth + ".tmp", synth)
    return lib


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(PKG, "libtdvh.so"))
