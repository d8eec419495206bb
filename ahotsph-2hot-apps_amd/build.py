"""Build recipe for libpgrav.so (gfx950 HIP kernels + C++ host runtime).

    python ahotsph-2hot-apps_amd/build.py            # build in-tree
    python ahotsph-2hot-apps_amd/build.py --asm      # also dump gfx950 assembly
    python ahotsph-2hot-apps_amd/build.py --define NAME[=VALUE] --out PATH
                                       # a variant library elsewhere (A/B timing;
                                       # load it with PGRAV_LIBRARY=PATH)

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


def build(verbose=False, asm=False, jobs=None, defines=(), out=None):
    out = os.path.abspath(out) if out else LIB
    bdir = BUILD if out == LIB else os.path.join(BUILD, "variant_" + os.path.basename(out))
    os.makedirs(bdir, exist_ok=True)
    dflags = ["-D" + d for d in defines]
    hip, cpp = _sources()
    cmds, objs = [], []
    for src in hip:
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + dflags + ["-c", src, "-o", obj])
        if asm:
            cmds.append([HIPCC] + HIP_FLAGS + dflags +
                        ["--cuda-device-only", "-S", src, "-o",
                         os.path.join(bdir, os.path.basename(src) + ".s")])
    for src in cpp:
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    jobs = jobs or min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for rc, cmd, log in ex.map(_run, cmds):
            if verbose or rc:
                print(cmd)
                if log:
                    print(log)
            if rc:
                raise RuntimeError("compile failed: " + cmd)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    rc, cmd, log = _run([HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
                        + ["-pthread"])
    if rc:
        print(cmd)
        print(log)
        raise RuntimeError("link failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--asm", action="store_true")
    ap.add_argument("--define", action="append", default=[])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    print(build(verbose=a.v, asm=a.asm, defines=a.define, out=a.out))
