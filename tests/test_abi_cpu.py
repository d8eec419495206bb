"""Host-side checks of the C ABI that need no GPU: the library loads, exports
every function include/pgrav.h declares, the Python binding declares each of
them, and host-side validation reports errors instead of crashing."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]
LIB = os.path.join(ROOT, "ahotsph-2hot-apps_amd", "libpgrav.so")


def _ensure_built():
    if not os.path.exists(LIB):
        import importlib.util
        spec = importlib.util.spec_from_file_location(
            "pgrav_build", os.path.join(ROOT, "ahotsph-2hot-apps_amd", "build.py"))
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.build()


def declared_functions():
    names = []
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.append(name)
    return sorted(set(names))


def test_header_declares_the_reference_abi():
    names = declared_functions()
    for ref in ("CUDA_Init", "CUDA_Finalize", "vecwidthCUDA", "qallocCUDA", "WalkInitSrcCUDA",
                "WalkTerminateSrcCUDA", "WalkInitSinkCUDA", "WalkTerminateSinkCUDA",
                "grav_mn_CUDA", "grav_mns_CUDA", "grav_mnss_CUDA", "grav_qns_CUDA",
                "grav_qnss_CUDA"):  # nln/gravcuda.h:13-69
        assert ref in names


def test_library_exports_every_declared_symbol():
    _ensure_built()
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    _ensure_built()
    from pgrav import _lib
    import pgrav
    pgrav.load_all()  # registers every part of the ABI
    missing = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing
    L = _lib.lib()
    for n in declared_functions():
        assert hasattr(L, n)


def test_host_validation_errors_without_gpu():
    """Without a visible GPU every entry point records an error (error mode 1)
    rather than aborting; with one, bad arguments are rejected before any
    launch.  Run in a subprocess so that the HIP runtime state stays private."""
    _ensure_built()
    code = r"""
import sys, numpy as np
sys.path.insert(0, %r)
import pgrav
from pgrav._lib import lib, PgravError
L = lib()
try:
    pgrav.GravCuda()
    ok = True
except PgravError as e:
    ok = False
    print("ERR", e)
if ok:
    g = pgrav.GravCuda()
    b = np.zeros(14 * 8, np.float32)
    g.walk_init_sink(b, 14)
    q, _ = g.qalloc()
    for bad in (lambda: g.grav_mns("nope", 0, 8, [0, 1], 1, 1.0, 1.0, q),
                lambda: g.grav_mn("pH", 4, 8, np.zeros(29 * 8, np.float32), 8, 29, 0.0, q),
                lambda: g.grav_mn("pH", 0, 8, np.zeros(29 * 8, np.float32), 7, 29, 0.0, q),
                lambda: g.grav_mn("pH", 0, 8, np.zeros(13 * 8, np.float32), 8, 13, 0.0, q)):
        try:
            bad()
            print("NOERR")
        except PgravError as e:
            print("ERR", e)
    g.walk_terminate_sink(b, 14)
print("DONE")
""" % os.path.join(ROOT, "ahotsph-2hot-apps_amd")
    r = subprocess.run(["python", "-c", code], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    lines = r.stdout.splitlines()
    assert lines[-1] == "DONE"
    assert "NOERR" not in r.stdout, r.stdout
    assert any(l.startswith("ERR") for l in lines)
