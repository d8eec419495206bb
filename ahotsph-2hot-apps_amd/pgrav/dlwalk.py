"""The reference's list producer for the drop-in path (pgrav_dl_* in include/pgrav.h).

    W = DLWalk(pos, mass, order=4, mac="tolerance", errtol=1e-4, frac_tol=0.01)
    W.calls      # (ncall, 6): routine (0 pH, 1 pQ), first slot, m, source offset, source_n, sz
    W.sources    # every SoA-8 block the calls, the pQ1 flushes and the leaves read
    W.tails      # (n, 6): each body's leaf remainders (H, Q, M offset / records)

With dl=dict(subtract_background=1) (periodic="images") the walk is DLRcritMACsb
(nln/mac.c:1565-1771): W.mflush / W.mss_* / W.mseg are the pMM_mnss_sK1 queue's
flushes, W.mmtails each leaf's MMvec remainder, W.background the near cube's field.

The walk is the reference's DL traversal with DLRcritMAC (nln/mac.c:1773-1844),
the sink inheritance of InheritSinkNlogN and the M x N batching of mxn_hexa /
mxn_quad2 (nln/mac.c:394-475, :477-1272); bodies are in the Morton-sorted order
of `perm` (slot i holds caller body perm[i]).
"""
import ctypes

import numpy as np

from ._lib import check, lib, register
from .batch import Params, make_params


class DLParams(ctypes.Structure):
    _fields_ = [("dlfac", ctypes.c_float), ("leaf_max_n", ctypes.c_int), ("gated", ctypes.c_int),
                ("hblock", ctypes.c_int), ("min_qsink", ctypes.c_int), ("min_hsink", ctypes.c_int),
                ("min_qsrc", ctypes.c_int), ("min_hsrc", ctypes.c_int), ("do_pQ", ctypes.c_int),
                ("do_pQL", ctypes.c_int), ("do_pH", ctypes.c_int), ("log_calls", ctypes.c_int),
                ("subtract_background", ctypes.c_int), ("do_pM", ctypes.c_int),
                ("min_msink", ctypes.c_int), ("min_msrc", ctypes.c_int), ("nx", ctypes.c_float),
                ("geometric_center", ctypes.c_int), ("tail_spans", ctypes.c_int)]


_vp, _i, _i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
register({
    "pgrav_dl_params_default": (None, [ctypes.POINTER(DLParams)]),
    "pgrav_dl_walk": (_vp, [_vp, _vp, _i64, ctypes.POINTER(Params), ctypes.POINTER(DLParams)]),
    "pgrav_dl_destroy": (None, [_vp]),
    "pgrav_dl_stats": (_i, [_vp, _vp]),
    "pgrav_dl_array": (_i64, [_vp, _i, _vp]),
})

# name -> (array id, dtype, values per element)
ARRAYS = {
    "cells": (0, np.float32, 8), "records": (1, np.float32, 29), "bodies": (2, np.float32, 4),
    "perm": (3, np.int64, 1), "child_off": (4, np.int64, 1), "children": (5, np.int32, 1),
    "calls": (6, np.int64, 6), "sources": (7, np.float32, 1), "qflush": (8, np.int64, 6),
    "qseg": (9, np.uint32, 2), "tails": (10, np.int64, 6), "interactions": (11, np.int64, 1),
    "log_calls": (12, np.int64, 4), "log_sources": (13, np.int32, 3),
    "log_appends": (14, np.int32, 3), "node_info": (15, np.int64, 4),
    # subtract_background
    "node_geom": (16, np.float32, 5), "ubs": (17, np.float32, 4), "body_leaf": (18, np.int64, 1),
    "mflush": (19, np.int64, 8), "mss_index": (20, np.uint16, 1), "mss_seg": (21, np.uint32, 2),
    "mss_source_n": (22, np.int32, 1), "mseg": (23, np.uint32, 2), "mcalls": (24, np.int64, 6),
    "mmtails": (25, np.int64, 3), "background": (26, np.float64, 3), "sb_sink": (27, np.float32, 6),
    "log_sb": (28, np.float32, 8), "log_flags": (29, np.int32, 2), "sb_stats": (30, np.int64, 1),
    # tail_spans
    "tail_runs": (31, np.int64, 3), "tail_index": (32, np.int64, 2),
}
SB_STATS = ("mns_flushes", "host_mono", "mm_segments", "anti_body", "anti_cell", "anti_empty",
            "mns_pairs")
STATS = ("mac_calls", "ph_calls", "pq_calls", "pq1_flushes", "source_floats", "mxn_calls",
         "leaf_h", "leaf_q", "leaf_m", "mxn_h_pairs", "mxn_q_pairs", "bad_interactions", "n",
         "nnodes", "nimage", "logged_calls")
ROUTINES = ("pH", "pQ")
SZ = {"pH": 29, "pQ": 13, "pM": 4}


def dl_params(**kw):
    d = DLParams()
    lib().pgrav_dl_params_default(ctypes.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise TypeError("unknown DL parameter " + k)
        setattr(d, k, v)
    return d


class DLWalk:
    def __init__(self, pos, mass, dl=None, **kw):
        """dl: dict of pgrav_dl_params fields (dlfac, leaf_max_n, gated, the M x N limits,
        log_calls); kw: pgrav_params (make_params)."""
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (pos.shape[0],)))
        self.params = make_params(**kw)
        self.dl = dl_params(**(dl or {}))
        self.h = lib().pgrav_dl_walk(pos.ctypes.data, mass.ctypes.data, pos.shape[0],
                                     ctypes.byref(self.params), ctypes.byref(self.dl))
        check()
        self._cache = {}

    def close(self):
        if getattr(self, "h", None):
            lib().pgrav_dl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        out = np.zeros(16, np.int64)
        lib().pgrav_dl_stats(self.h, out.ctypes.data)
        check()
        return dict(zip(STATS, (int(v) for v in out)))

    def array(self, name):
        if name not in self._cache:
            which, dt, per = ARRAYS[name]
            n = lib().pgrav_dl_array(self.h, which, None)
            check()
            a = np.empty(n * per, dt)
            if a.size:
                lib().pgrav_dl_array(self.h, which, a.ctypes.data)
                check()
            self._cache[name] = a.reshape(n, per) if per > 1 else a
        return self._cache[name]

    def __getattr__(self, name):
        if name in ARRAYS:
            return self.array(name)
        raise AttributeError(name)

    def block(self, off, nrec, sz):
        """nrec records of sz floats as SoA-8 blocks at source offset off (flat)."""
        return self.sources[off:off + nrec * sz]

    def leaf_runs(self, j):
        """Body slot j's leaf remainders as {list: [(source offset, records)]} (list 0
        Hvec, 1 Qvec, 2 Mvec), whichever way the walk emitted them."""
        out = {0: [], 1: [], 2: []}
        if self.dl.tail_spans:
            first, cnt = (int(v) for v in self.tail_index[j])
            for kind, off, nrec in self.tail_runs[first:first + cnt]:
                out[int(kind)].append((int(off), int(nrec)))
        else:
            t = self.tails[j]
            for k in range(3):
                if t[2 * k + 1]:
                    out[k].append((int(t[2 * k]), int(t[2 * k + 1])))
        return out

    def sb_stats(self):
        return dict(zip(SB_STATS, (int(v) for v in self.sb_stats_raw())))

    def sb_stats_raw(self):
        return self.array("sb_stats")

    def mac_calls(self):
        """The logged MAC calls: [(sink entity, [(entity, image, result)], [(list, entity,
        image)])] in call order (log_calls=1)."""
        lc, ls, la = self.log_calls, self.log_sources, self.log_appends
        out = []
        for k in range(lc.shape[0]):
            sink, s0, n, a0 = (int(v) for v in lc[k])
            a1 = int(lc[k + 1, 3]) if k + 1 < lc.shape[0] else la.shape[0]
            out.append((sink, ls[s0:s0 + n], la[a0:a1]))
        return out
