"""pgrav -- Python surface of the MI355X gravity interaction-list evaluation.

Mirrors nln/py_pgrav/pgrav.pyx (``monopole``, ``vec_monopole``: M x N
interactions on 1-D float32 arrays) and exposes the drop-in C ABI of
nln/gravcuda.h (class ``GravCuda``).  Every call runs the HIP kernels of
libpgrav.so; there is no CPU code path.
"""
import ctypes

import numpy as np

from ._lib import PgravError, Segment, check, lib, ptr  # noqa: F401

NSSE = 8  # SoA block width of the reference's AVX build (nln/vec.h:7)
MSZ, QSZ, HSZ = 4, 13, 29  # floats per order-0/2/4 record (nln/pgrav.cu:20, 33, 48)


def _f32(a, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or a.ndim != 1 \
            or not a.flags.c_contiguous:
        raise TypeError("%s must be a 1-D C-contiguous float32 array" % name)
    return a


def _mn(kind, p, accp, stride, f, sz, layout, e=0.0):
    _f32(p, "p")
    _f32(accp, "accp")
    _f32(f, "f")
    if stride < 4:
        raise ValueError("stride must be >= 4")
    m = p.shape[0] // stride
    if accp.shape[0] < m * stride:
        raise ValueError("accp is shorter than p")
    if layout == 0:
        n = f.shape[0] // sz
    else:
        n = f.shape[0] // sz
        if f.shape[0] % (sz * NSSE):
            raise ValueError("SoA-%d source array must hold whole blocks" % NSSE)
    lib().pgrav_mn_host(kind, ptr(p), ptr(accp), m, stride, ptr(f), n, sz, layout, e)
    check()


def monopole(p, accp, stride, f):
    """nln/py_pgrav/pgrav.pyx:21-32: m sinks (position at p[i*stride+1..3]) against
    n = len(f)/4 AoS monopoles (mass, x, y, z); accp[i*stride+0..3] += (ax, ay, az, phi)."""
    _mn(0, p, accp, stride, f, MSZ, 0)


def vec_monopole(p, accp, stride, f):
    """nln/py_pgrav/pgrav.pyx:36-47: as monopole() with f in SoA-NSSE blocks
    (nln/py_pgrav/test_pgrav.py:30 swizzle)."""
    _mn(0, p, accp, stride, f, MSZ, 1)


def mn(kind, p, accp, stride, f, layout=1, eps_inv=0.0):
    """Generic M x N: kind 'pM', 'pQ', 'pH' or 'pM_sK1' (nln/pgrav.cu:49-468) with
    SoA-8 (layout=1) or AoS (layout=0) records of 4/13/29 floats."""
    k = {"pM": (0, MSZ), "pQ": (2, QSZ), "pH": (4, HSZ), "pM_sK1": (5, MSZ)}[kind]
    _mn(k[0], p, accp, stride, f, k[1], layout, eps_inv)


def load_all():
    """Imports every submodule so that each part of the C ABI is declared."""
    import importlib
    import pkgutil
    for m in pkgutil.iter_modules(__path__):
        importlib.import_module(__name__ + "." + m.name)


def set_quad_phi_sign(sign):
    """+1: physical quadrupole potential (default); -1: nln/pgrav.cu's sign."""
    lib().pgrav_set_quad_phi_sign(int(sign))


def set_dropin_deferral(enable):
    """1 (default): the per-sink-group drop-in calls are batched and evaluated by
    one launch per kind; 0: one copy and launch per call.  Returns the old value."""
    return lib().pgrav_set_dropin_deferral(int(bool(enable)))


class GravCuda:
    """Thin object view of the drop-in ABI (nln/gravcuda.h) used by tests and
    by host drivers written in Python.  Arrays are kept alive by the caller."""

    def __init__(self):
        self.L = lib()
        self.L.CUDA_Init(None)
        check()

    def _call(self, name, *args):
        r = getattr(self.L, name)(*args)
        check()
        return r

    def vecwidth(self):
        return self._call("vecwidthCUDA")

    def qalloc(self, count=True):
        """(slot, busy slots); count=False passes NULL as nln/mac.c does (busy: -1)."""
        if not count:
            return self._call("qallocCUDA", None), -1
        inuse = ctypes.c_int(0)
        q = self._call("qallocCUDA", ctypes.cast(ctypes.byref(inuse), ctypes.c_void_p))
        return q, inuse.value

    def walk_init_sink(self, btab, stride):
        self.btab = btab
        self.stride = stride
        self._call("WalkInitSinkCUDA", ptr(btab), stride, btab.size // stride)

    def walk_terminate_sink(self, btab, stride):
        self._call("WalkTerminateSinkCUDA", ptr(btab), stride, btab.size // stride)

    def walk_init_src(self, qtab_bytes, stride_bytes, ncells):
        self._call("WalkInitSrcCUDA", ptr(qtab_bytes), stride_bytes, ncells)

    def walk_terminate_src(self):
        self._call("WalkTerminateSrcCUDA")

    def grav_mn(self, routine, first_body, m, f, source_n, sz, e, q):
        p = ctypes.c_void_p(self.btab.ctypes.data + 4 * first_body * self.stride)
        self._call("grav_mn_CUDA", routine.encode(), p, None, m, self.stride, ptr(f), source_n,
                   sz, e, None, q)

    def pinteract(self, first_body, m, f, source_n, sz, q):
        p = ctypes.c_void_p(self.btab.ctypes.data + 4 * first_body * self.stride)
        self._call("pinteractCUDA", p, None, m, self.stride, ptr(f), source_n, sz, q)

    def grav_mns(self, routine, base, m, seg, source_n, mmass, e, q):
        seg = np.ascontiguousarray(seg, np.uint32)
        self._call("grav_mns_CUDA", routine.encode(), base, m, ptr(seg), seg.size // 2,
                   source_n, mmass, e, None, q)

    def grav_mnss(self, routine, sink_base, m, ss_index, ss_seg, seg, source_base, source_n,
                  mmass, e, q):
        ss_index = np.ascontiguousarray(ss_index, np.uint16)
        ss_seg = np.ascontiguousarray(ss_seg, np.uint32)
        seg = np.ascontiguousarray(seg, np.uint32)
        source_n = np.ascontiguousarray(source_n, np.int32)
        self._call("grav_mnss_CUDA", routine.encode(), sink_base, m, ptr(ss_index), ptr(ss_seg),
                   ss_seg.size // 2, ptr(seg), seg.size // 2, source_base, ptr(source_n), mmass,
                   e, None, q)

    def grav_qns(self, routine, base, m, source_list, q):
        sl = np.ascontiguousarray(source_list, np.int32)
        self._call("grav_qns_CUDA", routine.encode(), base, m, ptr(sl), sl.size, q)

    def grav_qnss(self, routine, base, m, seg, source, source_n, q):
        seg = np.ascontiguousarray(seg, np.uint32)
        self._call("grav_qnss_CUDA", routine.encode(), base, m, ptr(seg), ptr(source),
                   source_n, q)

    def finalize(self):
        self._call("CUDA_Finalize")
