"""TEST INFRASTRUCTURE ONLY -- CPU baseline leg of bench.py.

Times the reference's SSE4 (4-wide) and AVX (8-wide) CPU kernels, compiled
from the reference sources into oracle/_ref, on a bounded sample of sink
groups of the benchmark workload, over exactly the interaction lists the GPU
evaluates: do_gravph_{sse4,avx8} (Plummer-softened order-4 cells, no dipole,
nln/do_grav_sse4.c:1768) or do_gravpq_{sse4,avx8} (order 2, :2090) over the
group's cells and do_gravp_{sse4,avx8}
(nln/do_grav_sse4.c:2167) over its bodies, per sink, as the leaf evaluation
of nln/mac.c:477-630 does.  Positions are absolute FP32, as in the reference.
"""
import ctypes
import os
import platform
import time

import numpy as np

from . import REF_DIR, lib, ref_available


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _run_lists(arch, D, groups, eps2, nthreads, cell="ph", csz=29):
    L = lib()
    fn = L.orc_cpu_ref_run_lists
    fn.restype = ctypes.c_double
    fn.argtypes = ([ctypes.c_char_p] * 3 + [ctypes.c_int] * 2 + [ctypes.c_void_p] * 10 +
                   [ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p])
    nsse = 4 if arch == "sse4" else 8
    out = np.zeros((D["sinks"].shape[0], 4), np.float32)
    ks = ctypes.c_double(0)
    groups = np.ascontiguousarray(groups, np.int32)
    path = os.path.join(REF_DIR, "libref_%s.so" % arch).encode()
    wall = fn(path, ("do_grav%s_%s" % (cell, arch)).encode(), ("do_gravp_" + arch).encode(), nsse, csz,
              D["rec"].ctypes.data, D["sinks"].ctypes.data, D["bodies"].ctypes.data,
              D["first"].ctypes.data, D["count"].ctypes.data, D["hoff"].ctypes.data,
              D["hidx"].ctypes.data, D["boff"].ctypes.data, D["bidx"].ctypes.data,
              groups.ctypes.data, groups.size, eps2, nthreads, out.ctypes.data, ctypes.byref(ks))
    if wall < 0:
        raise RuntimeError("reference CPU kernels could not be loaded from %s" % path)
    return wall, ks.value, out


def _interactions(D, groups):
    c = D["count"][groups].astype(np.int64)
    nh = D["hoff"][groups + 1] - D["hoff"][groups]
    nb = D["boff"][groups + 1] - D["boff"][groups]
    return int((c * (nh + nb)).sum())


def run_cpu_baseline(pos, mass, args, acc_gpu=None, target_s=15.0, nthreads=None, dec=None):
    """Returns the bench.py cpu_baseline object (interactions/s on the host cores).

    The sample is every group when the two timed passes fit in about
    `target_s` seconds of wall time, else a seeded random subset sized to fit.
    The rate is interactions / (kernel-call seconds summed over threads /
    threads): the gathers into SoA blocks are not counted against the CPU."""
    if not ref_available():
        raise RuntimeError("oracle/_ref is not built (make -C oracle ref)")
    from pgrav import batch
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    dec = dec or {"merge_groups": args.group_mode, "group_max": 64, "leaf_max": 16}
    H = batch.HostPlan(pos, mass, theta=args.theta, order=args.order, eps=args.eps,
                       soft="plummer", merge_groups=dec["merge_groups"],
                       group_max=dec["group_max"], leaf_max=dec["leaf_max"])
    st = H.stats()
    perm, spos = H.sorted()
    smass = np.asarray(mass, np.float64)[perm]
    # sinks in group order, so that every group is a run for the harness (bisection
    # groups are body sets); results come back in that order
    goff, gslots = H.group_slot_table()
    first, count = goff[:-1].copy(), np.diff(goff).astype(np.int32)
    # order-4 cells go to do_gravph (29-float records), order-2 cells to do_gravpq (their
    # first 13 floats: mass, centre, bmax, dipole, quadrupole); one cell order per plan
    kind, cell, csz = ("hexa", "ph", 29) if args.order >= 4 else ("quad", "pq", 13)
    hoff, hidx = H.csr(kind)
    boff, bidx = H.csr("body")
    others = [k for k in ("hexa", "quad", "mono") if k != kind]
    if any(H.csr(k)[1].size for k in others):
        raise RuntimeError("the CPU baseline covers one cell order plus bodies")
    D = {"rec": np.ascontiguousarray(H.records(np.arange(st["nnodes"], dtype=np.uint32))[:, :csz],
                                     np.float32),
         "sinks": np.ascontiguousarray(spos[gslots], np.float32),
         "bodies": np.ascontiguousarray(np.concatenate([smass[:, None], spos], axis=1), np.float32),
         "first": first, "count": count, "hoff": hoff, "hidx": hidx, "boff": boff, "bidx": bidx}
    del H
    ng = first.size
    rng = np.random.default_rng(777)
    eps2 = float(args.eps) ** 2
    result = {"unit": "interactions/s", "cores": nthreads, "kind": "reference",
              "cpu_model": _cpu_model()}
    # calibrate on a small sample, then size the timed sample to ~target_s seconds
    cal = rng.choice(ng, min(ng, 64 * nthreads), replace=False).astype(np.int32)
    wall, _, _ = _run_lists("sse4", D, cal, eps2, nthreads, cell, csz)
    per_int = wall / max(1, _interactions(D, cal))
    total = _interactions(D, np.arange(ng))
    # SSE4 then AVX (about half the time): ~1.5 SSE4 passes in the budget
    frac = min(1.0, target_s / 1.5 / max(per_int * total, 1e-9))
    groups = (np.arange(ng) if frac >= 1.0 else
              np.sort(rng.choice(ng, max(len(cal), int(frac * ng)), replace=False))).astype(np.int32)
    inter = _interactions(D, groups)
    for arch in ("sse4", "avx8"):
        # a workload that finishes in milliseconds is repeated (best pass kept)
        wall, ks, out = _run_lists(arch, D, groups, eps2, nthreads, cell, csz)
        for _ in range(min(50, int(1.0 / max(wall, 1e-4)))):
            w2, k2, _ = _run_lists(arch, D, groups, eps2, nthreads, cell, csz)
            if k2 < ks:
                wall, ks = w2, k2
        result[arch] = inter / (ks / nthreads)
        result[arch + "_wall_seconds"] = wall
        result[arch + "_kernel_seconds_per_thread"] = ks / nthreads
        if acc_gpu is not None and arch == "avx8":
            sidx = np.concatenate([np.arange(first[g], first[g] + count[g]) for g in groups[:4096]])
            ref = np.asarray(acc_gpu)[perm[gslots[sidx]]]
            d = (np.linalg.norm(out[sidx, :3] - ref[:, :3], axis=1) /
                 np.linalg.norm(ref[:, :3], axis=1))
            result["median_rel_diff_vs_gpu"] = float(np.median(d))
    result["value"] = max(result["sse4"], result["avx8"])
    result["sample"] = ("%s %d of %d sink groups (%d sinks, %.3e interactions) of the benchmark "
                        "workload: reference do_grav%s+do_gravp (SSE4 and AVX builds) per sink "
                        "over the same lists, %d threads, kernel calls timed"
                        % ("all" if groups.size == ng else "random", groups.size, ng,
                           int(count[groups].sum()), inter, cell, nthreads))
    result["prep_seconds"] = time.perf_counter() - t0
    return result
