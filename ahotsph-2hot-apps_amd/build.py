"""Build recipe for libpgrav.so (gfx950 HIP kernels + C++ host runtime).

    python ahotsph-2hot-apps_amd/build.py            # build in-tree
    python ahotsph-2hot-apps_amd/build.py --asm      # also dump gfx950 assembly

HIP sources are compiled by hipcc for gfx950 only; host-only C++ (tree, lists,
generators) by g++.  Everything links into one shared library next to this file
so that it travels with the source tree.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libpgrav.so")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-I" + INCLUDE, "-I" + CSRC,
             "-I" + HOST,
             # packed FP32 (v_pk_fma_f32 & co) stays off on this platform
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-march=x86-64-v2", "-I" + INCLUDE, "-I" + HOST,
             "-I" + CSRC,
             "-Wall", "-Wno-unused-function", "-pthread"]


def _sources():
    hip = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp")) \
        if os.path.isdir(HOST) else []
    return hip, cpp


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    return r.returncode, " ".join(cmd), out


def build(verbose=False, asm=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    hip, cpp = _sources()
    cmds, objs = [], []
    for src in hip:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        if asm:
            cmds.append([HIPCC] + HIP_FLAGS + ["--cuda-device-only", "-S", src, "-o",
                                               os.path.join(BUILD, os.path.basename(src) + ".s")])
    for src in cpp:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for rc, cmd, out in ex.map(_run, cmds):
            if verbose or rc:
                print(cmd)
                if out:
                    print(out)
            if rc:
                raise RuntimeError("compile failed: " + cmd)
    tmp = LIB + ".tmp"
    rc, cmd, out = _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
                        + ["-pthread"])
    if rc:
        print(cmd)
        print(out)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, asm="--asm" in sys.argv))
