"""Batched, mixed-precision force evaluation (part 3 of include/pgrav.h).

    plan = Plan(pos, mass, theta=0.6, order=4, eps=1e-4)
    plan.evaluate()                  # one launch over every sink group
    acc = plan.result()              # (n, 4): ax, ay, az, phi in the caller's order

The tree, moments and lists are built on the host inside Plan(); the GPU only
evaluates the lists.  Variants: REL32C (FP32 group-relative, compensated;
default), REL32, ABS32 (FP32 absolute, the reference's arithmetic), F64
(batched FP64) and PLAIN64 (FP64, one launch per group).
"""
import ctypes
import warnings

import numpy as np

from ._lib import PgravError, check, lib, register

REL32C, REL32, ABS32, F64, PLAIN64 = range(5)
VARIANTS = {"rel32c": REL32C, "rel32": REL32, "abs32": ABS32, "f64": F64, "plain64": PLAIN64}
SOFT = {"none": 0, "plummer": 1, "k1": 4}
KINDS = ("hexa", "quad", "mono", "body")


class Params(ctypes.Structure):
    _fields_ = [("theta", ctypes.c_double), ("group_max", ctypes.c_int),
                ("leaf_max", ctypes.c_int), ("order", ctypes.c_int), ("pcut", ctypes.c_int),
                ("soft_type", ctypes.c_int), ("eps", ctypes.c_double), ("G", ctypes.c_double),
                ("periodic", ctypes.c_int), ("box", ctypes.c_double), ("nthreads", ctypes.c_int),
                ("merge_groups", ctypes.c_int), ("persistent_blocks", ctypes.c_int),
                ("fp64", ctypes.c_int), ("dispatch", ctypes.c_int), ("mixed_order", ctypes.c_int),
                ("nimage", ctypes.c_int), ("rcb_cell", ctypes.c_int), ("mac", ctypes.c_int),
                ("errtol", ctypes.c_double), ("frac_tol", ctypes.c_double),
                ("frac_tol0", ctypes.c_double), ("ptol_boost", ctypes.c_double),
                ("stol_max", ctypes.c_double), ("quad_ncut", ctypes.c_int),
                ("hexa_ncut", ctypes.c_int), ("body_unroll", ctypes.c_int),
                ("ewald_ratio", ctypes.c_double), ("group_box", ctypes.c_int)]

# acceptance tests: opening angle theta; error tolerance (the reference's AREL_MAC
# critical radii, nln/cofm.c:276-565, taken lowest order first, nln/mac.c:1776-1800)
MAC = {"theta": 0, "tolerance": 1, "arel": 1, 0: 0, 1: 1}

# periodic boundary handling: none; nearest image + Ewald correction per pair
# (nln/ewald.c); explicit (2 nimage + 1)^3 images + local expansion of the far
# images (nln/mac.c:91-111, nln/main.c:1811-1862)
PERIODIC = {False: 0, True: 1, "none": 0, "ewald": 1, "images": 2, 0: 0, 1: 1, 2: 2}


class Stats(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnodes", ctypes.c_int64), ("ngroups", ctypes.c_int64),
                ("pairs", ctypes.c_int64 * 4), ("entries", ctypes.c_int64 * 4),
                ("t_tree", ctypes.c_double), ("t_lists", ctypes.c_double),
                ("t_upload", ctypes.c_double), ("mean_group", ctypes.c_double)]


_vp, _i, _i64, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
_P = ctypes.POINTER(Params)
_S = ctypes.POINTER(Stats)
register({
    "pgrav_params_default": (None, [_P]),
    "pgrav_plan_create": (_vp, [_vp, _vp, _i64, _P]),
    "pgrav_plan_destroy": (None, [_vp]),
    "pgrav_plan_update": (_i, [_vp, _vp, _vp]),
    "pgrav_plan_update_device": (_i, [_vp, _vp, _vp, _vp]),
    "pgrav_plan_adjoint": (_i, [_vp, _vp, _vp]),
    "pgrav_plan_adjoint_device": (_i, [_vp, _vp, _i, _vp, _vp]),
    "pgrav_plan_result_device": (_i, [_vp, _i, _vp, _i, _vp]),
    "pgrav_plan_stats": (_i, [_vp, _S]),
    "pgrav_plan_evaluate": (_i, [_vp, _i, _vp]),
    "pgrav_plan_set_launch": (_i, [_vp, _i]),
    "pgrav_plan_set_unroll": (_i, [_vp, _i]),
    "pgrav_plan_set_tile": (_i, [_vp, _i]),
    "pgrav_plan_set_dispatch": (_i, [_vp, _i]),
    "pgrav_plan_cells": (_i, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "pgrav_plan_time": (_i, [_vp, _i, _i, _vp, ctypes.POINTER(_d)]),
    "pgrav_plan_graph_capture": (_i, [_vp, _i]),
    "pgrav_plan_graph_launch": (_i, [_vp, _vp]),
    "pgrav_plan_synchronize": (_i, [_vp]),
    "pgrav_plan_result": (_i, [_vp, _i, _vp]),
    "pgrav_plan_direct": (_i, [_vp, _vp, _i, _vp]),
    "pgrav_plan_device_buffers": (_i, [_vp, _vp, _vp, _vp]),
    "pgrav_plan_permutation": (_i, [_vp, _vp]),
    "pgrav_host_plan": (_vp, [_vp, _vp, _i64, _P]),
    "pgrav_host_destroy": (None, [_vp]),
    "pgrav_host_stats": (_i, [_vp, _S]),
    "pgrav_host_group": (_i, [_vp, _i64, _vp, _vp, _vp]),
    "pgrav_host_list": (_i64, [_vp, _i64, _i, _vp, _i64]),
    "pgrav_host_list_images": (_i64, [_vp, _i64, _i, _vp, _i64]),
    "pgrav_host_group_slots": (_i64, [_vp, _i64, _vp]),
    "pgrav_host_group_slot_table": (_i, [_vp, _vp, _vp]),
    "pgrav_host_groups": (_i, [_vp, _vp, _vp, _vp]),
    "pgrav_host_csr": (_i64, [_vp, _i, _vp, _vp]),
    "pgrav_host_cell": (_i, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pgrav_host_rcrit": (_i, [_vp, _vp]),
    "pgrav_host_arel_rcrit": (_i, [_d, _d, _d, _d, _d, _i64, _d, _vp, _vp]),
    "pgrav_host_record": (_i, [_vp, _i64, _vp]),
    "pgrav_host_records": (_i, [_vp, _vp, _i64, _vp]),
    "pgrav_host_sorted": (_i, [_vp, _vp, _vp]),
    "pgrav_host_lpt": (_i, [_vp, _vp]),
    "pgrav_host_ewald_le": (_i, [_vp, _vp, _i64, _d, _i, _d, _vp]),
    "pgrav_host_ewald_le_sb": (_i, [_vp, _vp, _i64, _d, _i, _d, _vp]),
    "pgrav_ewald_le_device": (_i, [_vp, _vp, _i64, _d, _i, _d, _i, _vp, _vp]),
    "pgrav_host_le_lattice": (_i, [_i, _i, ctypes.POINTER(_d)]),
})


def make_params(theta=0.6, order=4, eps=1e-4, soft="plummer", G=1.0, group_max=64, leaf_max=16,
                pcut=0, periodic=False, box=1.0, nthreads=0, merge_groups=True,
                persistent_blocks=-1, fp64=True, dispatch=2, mixed_order=False, nimage=2,
                rcb_cell=32768, mac="theta", errtol=0.0, frac_tol=0.0, frac_tol0=0.0,
                ptol_boost=0.0, stol_max=0.0, quad_ncut=7, hexa_ncut=20, group_box=0,
                body_unroll=0, ewald_ratio=0.05):
    """pgrav_params; mac="tolerance" selects the error-tolerance acceptance, with
    errtol the absolute acceleration error per interaction (G = 1 units), frac_tol
    and frac_tol0 the relative bound and floor, quad_ncut / hexa_ncut the body
    counts from which a cell may be taken at order 2 / 4 (theta, pcut and
    mixed_order are then unused)."""
    p = Params()
    lib().pgrav_params_default(ctypes.byref(p))
    p.theta, p.order, p.eps, p.G = theta, order, eps, G
    p.soft_type = SOFT[soft] if isinstance(soft, str) else int(soft)
    p.group_max, p.leaf_max, p.pcut = group_max, leaf_max, pcut
    if periodic not in PERIODIC:
        raise ValueError("periodic must be one of %s" % sorted(map(str, PERIODIC)))
    p.periodic, p.box, p.nthreads = PERIODIC[periodic], box, nthreads
    p.nimage = nimage
    p.rcb_cell = rcb_cell
    p.merge_groups, p.persistent_blocks, p.fp64 = int(merge_groups), persistent_blocks, int(fp64)
    p.dispatch = dispatch
    p.mixed_order = int(mixed_order)
    if mac not in MAC:
        raise ValueError("mac must be one of %s" % sorted(map(str, MAC)))
    p.mac = MAC[mac]
    p.errtol, p.frac_tol, p.frac_tol0 = errtol, frac_tol, frac_tol0
    p.ptol_boost, p.stol_max = ptol_boost, stol_max
    p.quad_ncut, p.hexa_ncut = quad_ncut, hexa_ncut
    p.group_box = int(group_box)
    p.body_unroll = int(body_unroll)
    p.ewald_ratio = float(ewald_ratio)
    return p


def arel_rcrit(m, bmax, B2, B4, B6, count, r0=1.0, **kw):
    """(rcrit, rcrit_q, rcrit_h) of one cell under the tolerances of make_params(**kw)."""
    out = np.empty(3, np.float64)
    p = make_params(**kw)
    lib().pgrav_host_arel_rcrit(m, bmax, B2, B4, B6, count, r0, ctypes.byref(p), out.ctypes.data)
    check()
    return out


def _as_inputs(pos, mass):
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (pos.shape[0],)))
    return pos, mass


def _stats_dict(st):
    d = {"n": st.n, "nnodes": st.nnodes, "ngroups": st.ngroups, "mean_group": st.mean_group,
         "t_tree": st.t_tree, "t_lists": st.t_lists, "t_upload": st.t_upload}
    for k, name in enumerate(KINDS):
        d["pairs_" + name] = st.pairs[k]
        d["entries_" + name] = st.entries[k]
    d["pairs_total"] = sum(st.pairs[k] for k in range(4))
    return d


class Plan:
    """One force evaluation's device-resident tree, lists and results."""

    def __init__(self, pos, mass, **kw):
        self.pos, self.mass = _as_inputs(pos, mass)
        self.params = make_params(**kw)
        if self.params.periodic == 2 and self.params.mac == 0 and self.params.theta > 0.45:
            # measured on the 256^3 lattice: 1.8e-1 of the rms force at theta 0.6,
            # the accuracy of the Ewald-table mode (periodic 1) at 0.6 needs <= 0.4 here
            warnings.warn("explicit images (periodic='images') with the opening angle need "
                          "theta <= 0.4 on near-uniform boxes; theta %.2f errs by ~1e-1 of "
                          "the rms force (use periodic='ewald', a smaller theta, or "
                          "mac='tolerance')" % self.params.theta, stacklevel=2)
        self.n = self.pos.shape[0]
        L = lib()
        self.h = L.pgrav_plan_create(self.pos.ctypes.data, self.mass.ctypes.data, self.n,
                                     ctypes.byref(self.params))
        check()
        if not self.h:
            raise PgravError("pgrav_plan_create failed")
        self.version = 0  # bumped by every update: which bodies the device holds

    def close(self):
        if getattr(self, "h", None):
            lib().pgrav_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        st = Stats()
        lib().pgrav_plan_stats(self.h, ctypes.byref(st))
        check()
        return _stats_dict(st)

    def update(self, pos, mass):
        """Move the bodies, keeping the tree topology, groups and lists."""
        self.pos, self.mass = _as_inputs(pos, mass)
        self.version += 1
        lib().pgrav_plan_update(self.h, self.pos.ctypes.data, self.mass.ctypes.data)
        check()

    def update_device(self, pos_ptr, mass_ptr, stream=None):
        """update() from device arrays (caller order, FP64: n x 3 and n), ready on
        `stream` (a hipStream_t handle; None: the plan's stream)."""
        self.version += 1
        lib().pgrav_plan_update_device(self.h, pos_ptr, mass_ptr, stream)
        check()

    def result_device(self, variant, ptr, fp64=False, stream=None):
        """result() into a device array (n x 4 float32, or float64 with fp64),
        queued on `stream` after the last evaluation."""
        lib().pgrav_plan_result_device(self.h, _v(variant), ptr, int(fp64), stream)
        check()

    def adjoint_device(self, w_ptr, grad_ptr, w_fp64=False, stream=None):
        """adjoint() on device arrays: weights n x 4 (float32, or float64 with
        w_fp64), gradient n x 4 float64, both in caller order, on `stream`."""
        lib().pgrav_plan_adjoint_device(self.h, w_ptr, int(w_fp64), grad_ptr, stream)
        check()

    def evaluate(self, variant=REL32C, stream=None):
        lib().pgrav_plan_evaluate(self.h, _v(variant), stream)
        check()

    def adjoint(self, w4):
        """Gradient of L = sum_i w_i.a_i + v_i phi_i (w4 = (w, v) per body, caller
        order) with respect to positions and masses: (n, 4) = (dL/dx, dL/dm)."""
        w4 = np.ascontiguousarray(w4, np.float64).reshape(self.n, 4)
        out = np.empty((self.n, 4), np.float64)
        lib().pgrav_plan_adjoint(self.h, w4.ctypes.data, out.ctypes.data)
        check()
        return out

    def set_launch(self, persistent_blocks):
        lib().pgrav_plan_set_launch(self.h, int(persistent_blocks))
        check()

    def set_unroll(self, unroll):
        lib().pgrav_plan_set_unroll(self.h, int(unroll))
        check()

    def set_tile(self, tile):
        """List entries per LDS round of the rel32c kernels: 32, 64 or 128."""
        lib().pgrav_plan_set_tile(self.h, int(tile))
        check()

    def set_dispatch(self, mode):
        lib().pgrav_plan_set_dispatch(self.h, int(mode))
        check()

    def time(self, variant=REL32C, reps=3, stream=None):
        ms = ctypes.c_double(0)
        lib().pgrav_plan_time(self.h, _v(variant), reps, stream, ctypes.byref(ms))
        check()
        return ms.value

    def graph_capture(self, variant=REL32C):
        lib().pgrav_plan_graph_capture(self.h, _v(variant))
        check()

    def graph_launch(self, stream=None):
        lib().pgrav_plan_graph_launch(self.h, stream)
        check()

    def synchronize(self):
        lib().pgrav_plan_synchronize(self.h)
        check()

    def result(self, variant=REL32C):
        out = np.empty((self.n, 4), np.float64)
        lib().pgrav_plan_result(self.h, _v(variant), out.ctypes.data)
        check()
        return out

    def direct(self, sample):
        """FP64 direct sum (same softening and G) for bodies `sample` (original indices)."""
        s = np.ascontiguousarray(sample, np.int64)
        out = np.empty((s.size, 4), np.float64)
        lib().pgrav_plan_direct(self.h, s.ctypes.data, s.size, out.ctypes.data)
        check()
        return out

    def device_buffers(self):
        o32, o64, bp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        lib().pgrav_plan_device_buffers(self.h, ctypes.byref(o32), ctypes.byref(o64),
                                        ctypes.byref(bp))
        check()
        return o32.value, o64.value, bp.value

    def cells(self, node0=0, count=None):
        """Device FP64 (cofm4, bmax, raw35) of nodes node0 .. node0+count."""
        count = self.stats()["nnodes"] - node0 if count is None else count
        c4 = np.empty((count, 4), np.float64)
        bm = np.empty(count, np.float64)
        raw = np.empty((count, 35), np.float64)
        lib().pgrav_plan_cells(self.h, node0, count, c4.ctypes.data, bm.ctypes.data,
                               raw.ctypes.data)
        check()
        return c4, bm, raw

    def permutation(self):
        perm = np.empty(self.n, np.int64)
        lib().pgrav_plan_permutation(self.h, perm.ctypes.data)
        check()
        return perm


def _v(variant):
    return VARIANTS[variant] if isinstance(variant, str) else int(variant)


class HostPlan:
    """The host half alone: tree, groups and lists, inspectable without a GPU."""

    def __init__(self, pos, mass, **kw):
        self.pos, self.mass = _as_inputs(pos, mass)
        self.params = make_params(**kw)
        if self.params.periodic == 2 and self.params.mac == 0 and self.params.theta > 0.45:
            # measured on the 256^3 lattice: 1.8e-1 of the rms force at theta 0.6,
            # the accuracy of the Ewald-table mode (periodic 1) at 0.6 needs <= 0.4 here
            warnings.warn("explicit images (periodic='images') with the opening angle need "
                          "theta <= 0.4 on near-uniform boxes; theta %.2f errs by ~1e-1 of "
                          "the rms force (use periodic='ewald', a smaller theta, or "
                          "mac='tolerance')" % self.params.theta, stacklevel=2)
        self.n = self.pos.shape[0]
        self.h = lib().pgrav_host_plan(self.pos.ctypes.data, self.mass.ctypes.data, self.n,
                                       ctypes.byref(self.params))
        check()

    def __del__(self):
        if getattr(self, "h", None):
            lib().pgrav_host_destroy(self.h)
            self.h = None

    def stats(self):
        st = Stats()
        lib().pgrav_host_stats(self.h, ctypes.byref(st))
        check()
        return _stats_dict(st)

    def group(self, g):
        first, cnt = ctypes.c_int64(0), ctypes.c_int(0)
        c = np.empty(4, np.float64)
        lib().pgrav_host_group(self.h, g, ctypes.byref(first), ctypes.byref(cnt), c.ctypes.data)
        check()
        return first.value, cnt.value, c

    def list(self, g, kind):
        k = KINDS.index(kind) if isinstance(kind, str) else kind
        n = lib().pgrav_host_list(self.h, g, k, None, 0)
        check()
        out = np.empty(n, np.uint32)
        lib().pgrav_host_list(self.h, g, k, out.ctypes.data, n)
        return out

    def group_slots(self, g):
        """Sorted body slots of group g (a run, or the body set of a bisection group)."""
        n = lib().pgrav_host_group_slots(self.h, g, None)
        check()
        out = np.empty(n, np.int64)
        lib().pgrav_host_group_slots(self.h, g, out.ctypes.data)
        return out

    def group_slot_table(self):
        """(off[G+1], slots[N]): every group's sink slots, in group order."""
        st = self.stats()
        off = np.empty(st["ngroups"] + 1, np.int64)
        slots = np.empty(st["n"], np.int64)
        lib().pgrav_host_group_slot_table(self.h, off.ctypes.data, slots.ctypes.data)
        check()
        return off, slots

    def list_images(self, g, kind):
        """periodic = 2: the image of each entry of list(g, kind), numbered
        (i + n)(2n + 1)^2 + (j + n)(2n + 1) + (k + n) for the offset (i, j, k) box."""
        k = KINDS.index(kind) if isinstance(kind, str) else kind
        n = lib().pgrav_host_list_images(self.h, g, k, None, 0)
        check()
        out = np.empty(n, np.uint8)
        lib().pgrav_host_list_images(self.h, g, k, out.ctypes.data, n)
        return out

    def cell(self, node):
        c = np.empty(3, np.float64)
        raw = np.empty(35, np.float64)
        bmax, first = ctypes.c_double(0), ctypes.c_int64(0)
        cnt, child, nch = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        lib().pgrav_host_cell(self.h, node, c.ctypes.data, ctypes.byref(bmax), raw.ctypes.data,
                              ctypes.byref(first), ctypes.byref(cnt), ctypes.byref(child),
                              ctypes.byref(nch))
        check()
        return {"cofm": c, "bmax": bmax.value, "raw": raw, "first": first.value,
                "count": cnt.value, "child": child.value, "nchild": nch.value}

    def rcrit(self):
        """(nnodes, 3) critical radii of orders 0, 2, 4 (tolerance-MAC plans)."""
        out = np.empty((self.stats()["nnodes"], 3), np.float64)
        lib().pgrav_host_rcrit(self.h, out.ctypes.data)
        check()
        return out

    def record(self, node):
        rec = np.empty(29, np.float64)
        lib().pgrav_host_record(self.h, node, rec.ctypes.data)
        check()
        return rec

    def groups(self):
        """(first, count, centre4) of every group."""
        G = self.stats()["ngroups"]
        first = np.empty(G, np.int64)
        cnt = np.empty(G, np.int32)
        cen = np.empty((G, 4), np.float64)
        lib().pgrav_host_groups(self.h, first.ctypes.data, cnt.ctypes.data, cen.ctypes.data)
        check()
        return first, cnt, cen

    def csr(self, kind):
        """One list kind for all groups: (off[G+1], idx)."""
        k = KINDS.index(kind) if isinstance(kind, str) else kind
        G = self.stats()["ngroups"]
        off = np.empty(G + 1, np.int64)
        n = lib().pgrav_host_csr(self.h, k, off.ctypes.data, None)
        check()
        idx = np.empty(max(n, 0), np.uint32)
        lib().pgrav_host_csr(self.h, k, None, idx.ctypes.data)
        check()
        return off, idx

    def records(self, nodes):
        nodes = np.ascontiguousarray(nodes, np.uint32)
        rec = np.empty((nodes.size, 29), np.float64)
        lib().pgrav_host_records(self.h, nodes.ctypes.data, nodes.size, rec.ctypes.data)
        check()
        return rec

    def sorted(self):
        perm = np.empty(self.n, np.int64)
        pos = np.empty((self.n, 3), np.float64)
        lib().pgrav_host_sorted(self.h, perm.ctypes.data, pos.ctypes.data)
        check()
        return perm, pos

    def lpt(self):
        o = np.empty(self.stats()["ngroups"], np.int32)
        lib().pgrav_host_lpt(self.h, o.ctypes.data)
        check()
        return o


def ewald_le_device(pos_ptr, mass_ptr, n, acc_ptr, box=1.0, nimage=2, G=1.0,
                    subtract_background=False, stream=None):
    """The far field on the GPU for device arrays (FP64: n x 3 positions, n masses);
    adds into acc_ptr (n x 4 FP64), queued on `stream` (hipStream_t or None)."""
    rc = lib().pgrav_ewald_le_device(pos_ptr, mass_ptr, int(n), float(box), int(nimage), float(G),
                                     int(bool(subtract_background)), acc_ptr, stream)
    check()
    if rc:
        raise PgravError("pgrav_ewald_le_device failed")


def ewald_le_host(pos, mass, box=1.0, nimage=2, G=1.0, subtract_background=False):
    """Far images + background of the explicit-image scheme at each body, FP64 on
    the host: (n, 4) = (a, phi), what FixCubeEwaldLE adds (nln/main.c:1836-1860);
    subtract_background: its form for the background-subtracted lists."""
    pos, mass = _as_inputs(pos, mass)
    out = np.zeros((pos.shape[0], 4))
    fn = lib().pgrav_host_ewald_le_sb if subtract_background else lib().pgrav_host_ewald_le
    rc = fn(pos.ctypes.data, mass.ctypes.data, pos.shape[0], float(box), int(nimage), float(G),
            out.ctypes.data)
    check()
    if rc:
        raise PgravError("pgrav_host_ewald_le failed")
    return out


def le_lattice(k, nimage):
    """x^k coefficient c_k of the far images' sum of 1/|r - n| (k = 4, 6, 8)."""
    c = ctypes.c_double()
    rc = lib().pgrav_host_le_lattice(int(k), int(nimage), ctypes.byref(c))
    check()
    if rc:
        raise PgravError("pgrav_host_le_lattice failed")
    return c.value


def rel_errors(acc, ref):
    """Per-body relative acceleration and potential errors."""
    da = np.linalg.norm(acc[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    dp = np.abs(acc[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
    return da, dp
