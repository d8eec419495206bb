"""TEST INFRASTRUCTURE ONLY -- CPU baseline leg of bench.py.

Times the reference's SSE4 (4-wide) and AVX (8-wide) CPU kernels, compiled
from the reference sources into oracle/_ref, on a bounded sample of sink
groups of the benchmark workload, over exactly the interaction lists the GPU
evaluates: do_gravph_{sse4,avx8} (Plummer-softened order-4 cells, no dipole,
nln/do_grav_sse4.c:1768) over the group's cells and do_gravp_{sse4,avx8}
(nln/do_grav_sse4.c:2167) over its bodies, per sink, as the leaf evaluation
of nln/mac.c:477-630 does.  Positions are absolute FP32, as in the reference.
"""
import ctypes
import os
import platform
import time

import numpy as np

from . import HERE, REF_DIR, lib, ref_available, to_soa


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _prepare(H, groups, spos, smass, nsse):
    sink_off, cell_off, body_off = [0], [0], [0]
    sinks, cells, bodies = [], [], []
    inter = 0
    for g in groups:
        first, cnt, _ = H.group(int(g))
        sinks.append(spos[first:first + cnt].astype(np.float32))
        nodes = H.list(int(g), "hexa")
        recs = H.records(nodes)
        body = H.list(int(g), "body")
        b4 = np.concatenate([smass[body, None], spos[body]], axis=1) if body.size else \
            np.zeros((0, 4))
        cells.append(to_soa(recs.astype(np.float32), nsse) if len(recs) else np.zeros(0, np.float32))
        bodies.append(to_soa(b4.astype(np.float32), nsse) if body.size else np.zeros(0, np.float32))
        sink_off.append(sink_off[-1] + cnt)
        cell_off.append(cell_off[-1] + (len(recs) + nsse - 1) // nsse)
        body_off.append(body_off[-1] + (body.size + nsse - 1) // nsse)
        inter += cnt * (len(recs) + body.size)
    cat = lambda xs: np.ascontiguousarray(np.concatenate(xs) if xs else np.zeros(0), np.float32)
    return (np.array(sink_off, np.int64), cat([s.reshape(-1) for s in sinks]),
            np.array(cell_off, np.int64), cat(cells), np.array(body_off, np.int64), cat(bodies),
            inter)


def _run(arch, prep, ngroups, eps2, nthreads, reps):
    L = lib()
    fn = L.orc_cpu_ref_run
    fn.restype = ctypes.c_double
    fn.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                   ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_float, ctypes.c_int,
                                                            ctypes.c_int, ctypes.c_void_p]
    sink_off, sinks, cell_off, cells, body_off, bodies, inter = prep
    nsse = 4 if arch == "sse4" else 8
    out = np.zeros((sink_off[-1], 4), np.float32)
    path = os.path.join(REF_DIR, "libref_%s.so" % arch).encode()
    t = fn(path, ("do_gravph_" + arch).encode(), ("do_gravp_" + arch).encode(), nsse, 29, ngroups,
           sink_off.ctypes.data, sinks.ctypes.data, cell_off.ctypes.data, cells.ctypes.data,
           body_off.ctypes.data, bodies.ctypes.data, eps2, nthreads, reps, out.ctypes.data)
    if t < 0:
        raise RuntimeError("reference CPU kernels could not be loaded from %s" % path)
    return t, out


def run_cpu_baseline(pos, mass, args, acc_gpu=None, target_s=15.0, nthreads=None):
    """Returns the bench.py cpu_baseline object (interactions/s on the host cores)."""
    if not ref_available():
        raise RuntimeError("oracle/_ref is not built (make -C oracle ref)")
    from pgrav import batch
    nthreads = nthreads or min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    H = batch.HostPlan(pos, mass, theta=args.theta, order=args.order, eps=args.eps,
                       soft="plummer", merge_groups=args.group_mode)
    perm, spos = H.sorted()
    smass = np.asarray(mass, np.float64)[perm]
    ng = H.stats()["ngroups"]
    rng = np.random.default_rng(777)
    eps2 = float(args.eps) ** 2
    result = {"unit": "interactions/s", "cores": nthreads, "kind": "reference",
              "cpu_model": _cpu_model()}
    # calibrate on a small sample, then size the timed sample to ~target_s seconds
    k = min(ng, 16 * nthreads)
    groups = rng.choice(ng, k, replace=False)
    prep = _prepare(H, groups, spos, smass, 4)
    t, _ = _run("sse4", prep, k, eps2, nthreads, 1)
    rate = prep[-1] / t
    k2 = int(min(ng, max(k, k * target_s / 2 / max(t, 1e-3))))
    groups = rng.choice(ng, k2, replace=False)
    for arch in ("sse4", "avx8"):
        nsse = 4 if arch == "sse4" else 8
        prep = _prepare(H, groups, spos, smass, nsse)
        t, out = _run(arch, prep, k2, eps2, nthreads, 1)
        result[arch] = prep[-1] / t
        result[arch + "_seconds"] = t
        if acc_gpu is not None and arch == "sse4":
            # the same bodies' results from the GPU (caller order -> sorted order)
            sidx = np.concatenate([np.arange(*(lambda f, c: (f, f + c))(*H.group(int(g))[:2]))
                                   for g in groups])
            ref = np.asarray(acc_gpu)[perm[sidx]]
            d = np.linalg.norm(out[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
            result["median_rel_diff_vs_gpu"] = float(np.median(d))
    result["value"] = max(result["sse4"], result["avx8"])
    result["sample"] = ("%d of %d sink groups (%d sinks, %.3e interactions) of the benchmark "
                        "workload, do_gravph+do_gravp per sink, %d threads"
                        % (k2, ng, prep[0][-1], prep[-1], nthreads))
    result["prep_seconds"] = time.perf_counter() - t0
    return result
