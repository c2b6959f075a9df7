"""Builds nova_amd/lib/libnova.so: the drop-in C-ABI library (host C++ + HIP for gfx950).

Host translation units (*.cpp) are compiled with g++ -O3 -ffp-contract=off so the CPU entry
points round exactly like nova's own g++ build; device translation units (*.hip) with hipcc
for gfx950 only, with packed-FP32 instruction selection disabled. Objects go to
nova_amd/_build/, the shared library to nova_amd/lib/libnova.so.

Usage: python -m nova_amd.build  (or python nova_amd/build.py)
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libnova.so")
INCLUDE = os.path.join(ROOT, "include")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")

CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Wall", "-Wno-unused-function", "-I" + INCLUDE]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-ffp-contract=off", "-Wno-unused-result", "-I" + INCLUDE]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n  %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def sources():
    names = sorted(os.listdir(CSRC))
    return ([os.path.join(CSRC, f) for f in names if f.endswith(".cpp")],
            [os.path.join(CSRC, f) for f in names if f.endswith(".hip")])


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cpp, hip = sources()
    cmds = []
    objs = []
    for src in cpp:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([CXX] + CXX_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    for src in hip:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for c, o in zip(cmds, outs):
            print(" ".join(c))
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


# Host sanitizers (AddressSanitizer, UndefinedBehaviorSanitizer) on the host translation
# units: libnova_san.so links them, built with the sanitizers, to the device translation
# units' objects of the plain build (device code is never sanitized). For tests of the *_cpu
# entry points with the sanitizer runtime preloaded.
SAN_FLAGS = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined",
             "-fno-omit-frame-pointer", "-g"]
SAN_LIB = os.path.join(LIBDIR, "libnova_san.so")


def build_sanitized(jobs=None):
    cpp, hip = sources()
    hipObjs = [os.path.join(BUILD, os.path.basename(src) + ".o") for src in hip]
    if not all(os.path.exists(o) for o in hipObjs):
        build(jobs=jobs)
    sdir = os.path.join(BUILD, "san")
    os.makedirs(sdir, exist_ok=True)
    cmds, objs = [], []
    for src in cpp:
        obj = os.path.join(sdir, os.path.basename(src) + ".o")
        cmds.append([CXX] + CXX_FLAGS + SAN_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    host_san = []
    for f in SAN_FLAGS[:3]:
        host_san += ["-Xarch_host", f]
    tmp = SAN_LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH] + host_san + ["-o", tmp] +
         objs + hipObjs)
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        print(build(verbose="-v" in sys.argv))
