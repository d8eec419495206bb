"""Measurements of the path on the configurations it is specified on (one GPU).

    python studies.py c2 [--n 2097152]
    python studies.py c3 [--n 16777216] [--thetas 0.5 0.6 0.7] [--group-boxes 0 1]
                             [--variants rel32c rel32 abs32 f64]
    python studies.py c4 [--ng 256] [--theta 0.6 | --thetas 0.4 0.5 0.6]
                             [--scheme ewald | images --nimage 2]
                             [--mac tolerance --errtols 1e-2 1e-3] [--ewald-ratio 0.05]
    python studies.py c5 [--n 1048576]
    python studies.py dropin [--n 262144]
    python studies.py dropin_sb [--ng 32]
    python studies.py steps [--n 16777216]
    python studies.py mac [--n 16777216] [--errtols 1e-2 1e-3 1e-4] [--leaves 16 64]

c2  the plain method on the GPU at scale: clustered bodies (pgrav.ics.clustered),
    order 4, theta 0.6, Plummer eps 1e-4, FP64 with one launch per sink group
    of <= 64 bodies (the reference's dispatch, nln/pgrav.cu:1078), beside the
    batched FP64 kernel and the FP32 group-relative compensated kernel on the
    same lists; errors against the FP64 direct sum on 8192 sampled bodies.
c3  clustered bodies (pgrav.ics.clustered), order 4, Plummer eps 1e-4, for each
    opening angle: kernel time of the FP32 group-relative compensated (rel32c),
    plain-sum (rel32) and absolute-coordinate (abs32, the reference's arithmetic)
    variants and of the batched FP64 variant, each with its force error against
    the FP64 direct sum on 8192 sampled bodies and, for the FP32 variants, against
    the FP64 kernel on the same lists over all bodies (the FP32 rounding alone);
    --group-boxes 1 measures the acceptance distance from each group's bounding
    box instead of its sphere; --offset K shifts the system by K system sizes.
c4  periodic unit box, ng^3 lattice displaced by a Gaussian random field
    (pgrav.ics.lattice_grf): eager and graph-replayed evaluation time, error
    against the periodic FP64 direct sum (nearest image + 64^3 Ewald table) on
    1024 sampled bodies, net momentum.  --scheme ewald: nearest image + the
    Ewald correction per pair; images: (2 nimage + 1)^3 explicit images + the
    far field by the local expansion (the reference's FixCubeEwaldLE scheme).
    --mac tolerance: after the theta row, the error-tolerance acceptance with
    errtol = each of --errtols times the rms force of the sampled bodies.
c5  clustered N bodies, tree and lists fixed: forward and adjoint wall time, and
    the adjoint against central finite differences (h = 1e-5 and h/2, combined
    by Richardson extrapolation; FP64 forward) of L = sum_i w_i.a_i + v_i phi_i
    on 64 sampled coordinates.
dropin  the reference's own GPU dispatch through the drop-in ABI (nln/gravcuda.h):
    per sink group one slot (qallocCUDA), grav_mn_CUDA("pH") on the group's
    order-4 cells as SoA-8 Hvec blocks and grav_mns_CUDA("pMM_sK1") on its body
    segments (the calls of nln/mac.c:1218 and the leaf path), results added
    into the body table by WalkTerminateSinkCUDA; timed from the first call to
    the end of WalkTerminateSinkCUDA with the lists already built, against the
    batched kernel on the same lists (K1 bodies, unsoftened cells, as
    pgrav.cu).
dropin_sb  the subtract_background walk (pgrav_dl_walk, geometric centres, one
    image layer) on an ng^3 lattice (rms displacement 0.3 spacings): its pH / pQ
    calls, pQ1 flushes and pMM_mnss_sK1 flushes through the drop-in ABI, all and
    each class alone, timed from WalkInitSinkCUDA to WalkTerminateSinkCUDA.
steps  the moment update under a fixed tree and the kick-drift of the body
    table at the benchmark's N, inputs on the device, 5 calls each.
mac  clustered bodies as c3, the error-tolerance acceptance (mac = 1, the
    reference's AREL_MAC radii taken lowest order first) against the opening
    angle: errtol = each of --errtols times the median |a| of the sampled bodies,
    frac_tol 0, bisection groups, each leaf size of --leaves; kernel time of
    rel32c, list composition and the error against the FP64 direct sum on 8192
    sampled bodies; first a theta = 0.6 row with the tuned decomposition.

Each configuration prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import sys
import time

import os

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ahotsph-2hot-apps_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

from pgrav import autotune, batch, ics  # noqa: E402


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _errors(acc, ref):
    da, dp = batch.rel_errors(acc, ref)
    return {"median": float(np.median(da)), "p90": float(np.percentile(da, 90)),
            "p99": float(np.percentile(da, 99)), "phi_median": float(np.median(dp))}


def _decomposition(n, theta, workload, eps=1e-4):
    dec = {"merge_groups": 1, "group_max": 64, "leaf_max": 16, "persistent_blocks": -1,
           "body_unroll": 4, "dispatch": 2}
    t = autotune.load(autotune.workload_key(n, theta, 4, "plummer", eps, workload))
    if t:
        dec.update({k: t["best"][k] for k in dec if k in t["best"]})
    return dec


def c2(args):
    pos, mass = ics.clustered(args.n, seed=2)
    rng = np.random.default_rng(12345)
    sample = rng.choice(pos.shape[0], min(8192, pos.shape[0]), replace=False)
    dec = _decomposition(pos.shape[0], 0.6, "clustered")
    t0 = time.perf_counter()
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4, soft="plummer", fp64=True, **dec)
    st = plan.stats()
    _log("c2: plan %.1fs, %d groups (mean %.1f), %.3e pairs" %
         (time.perf_counter() - t0, st["ngroups"], st["mean_group"], st["pairs_total"]))
    ref = plan.direct(sample)
    row = {"config": "c2", "n": int(pos.shape[0]), "theta": 0.6, "decomposition": dec,
           "groups": st["ngroups"], "mean_group": st["mean_group"], "pairs": st["pairs_total"],
           "variants": {}}
    for var, reps in (("plain64", 1), ("f64", 3), ("rel32c", 3)):
        plan.evaluate(var)
        ms = plan.time(var, reps=reps)
        e = _errors(plan.result(var)[sample], ref)
        row["variants"][var] = {"ms": ms, "interactions_per_s": st["pairs_total"] / (ms * 1e-3),
                                "errors": e}
        _log("   %-8s %10.3f ms  %.3e interactions/s  median %.2e  p99 %.2e"
             % (var, ms, st["pairs_total"] / (ms * 1e-3), e["median"], e["p99"]))
    plan.close()
    print(json.dumps(row), flush=True)


def c3(args):
    pos, mass = ics.clustered(args.n, seed=2)
    # --offset moves the whole system by that many system sizes: the lists and the
    # FP64 results do not change, the absolute FP32 coordinates lose bits
    pos = pos + args.offset * (np.ptp(pos, axis=0).max())
    rng = np.random.default_rng(12345)
    sample = rng.choice(pos.shape[0], min(8192, pos.shape[0]), replace=False)
    ref = None
    for theta, gbox in [(t, b) for t in args.thetas for b in args.group_boxes]:
        dec = _decomposition(pos.shape[0], theta, "clustered")
        t0 = time.perf_counter()
        plan = batch.Plan(pos, mass, theta=theta, order=4, eps=1e-4, soft="plummer",
                          fp64="f64" in args.variants, group_box=gbox, **dec)
        st = plan.stats()
        _log("c3 theta %.2f group_box %d: plan %.1fs, %d groups, %.3e pairs" %
             (theta, gbox, time.perf_counter() - t0, st["ngroups"], st["pairs_total"]))
        if ref is None:  # the direct sum does not depend on theta
            ref = plan.direct(sample)
        row = {"config": "c3", "n": int(pos.shape[0]), "theta": theta, "group_box": gbox,
               "decomposition": dec, "groups": st["ngroups"], "pairs": st["pairs_total"],
               "pairs_hexa": st["pairs_hexa"], "pairs_body": st["pairs_body"], "variants": {}}
        # the batched FP64 kernel first: every FP32 variant is also compared with it on
        # all bodies (same lists, so what differs is the FP32 arithmetic alone)
        f64 = None
        for var in sorted(args.variants, key=lambda v: v != "f64"):
            plan.evaluate(var)
            ms = plan.time(var, reps=3)
            res = plan.result(var)
            e = _errors(res[sample], ref)
            row["variants"][var] = {"kernel_ms": ms, "interactions_per_s": st["pairs_total"] / (ms * 1e-3),
                                    "errors": e}
            if var == "f64":
                f64 = res.copy()
            elif f64 is not None:
                da, dp = batch.rel_errors(res, f64)
                row["variants"][var]["vs_f64_same_lists"] = {
                    "median": float(np.median(da)), "p99": float(np.percentile(da, 99)),
                    "max": float(da.max()), "phi_median": float(np.median(dp))}
                _log("   %-6s vs f64, all bodies: median %.2e  p99 %.2e  max %.2e" %
                     (var, np.median(da), np.percentile(da, 99), da.max()))
            _log("   %-6s %9.3f ms  median %.2e  p99 %.2e" % (var, ms, e["median"], e["p99"]))
        plan.close()
        print(json.dumps(row), flush=True)


def _c4_run(args, pos, mass, label, sample, ref, **kw):
    n = pos.shape[0]
    t0 = time.perf_counter()
    plan = batch.Plan(pos, mass, periodic=args.scheme, nimage=args.nimage, box=1.0, order=4,
                      eps=args.eps, soft="plummer", fp64=False, ewald_ratio=args.ewald_ratio,
                      **kw)
    st = plan.stats()
    _log("c4 %s %d^3: plan %.1fs, %d groups, %.3e pairs" % (label, args.ng, time.perf_counter() - t0,
                                                            st["ngroups"], st["pairs_total"]))
    plan.evaluate("rel32c")
    eager = plan.time("rel32c", reps=3)
    plan.graph_capture("rel32c")
    plan.graph_launch()
    plan.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.graph_launch()
    plan.synchronize()
    graph = (time.perf_counter() - t0) / reps * 1e3
    acc = plan.result("rel32c")
    if ref is None:
        ref = plan.direct(sample)
    err = np.linalg.norm(acc[sample, :3] - ref[:, :3], axis=1)
    rms = float(np.sqrt((ref[:, :3] ** 2).sum(1).mean()))
    net = np.abs((mass[:, None] * acc[:, :3]).sum(0)).max()
    typical = (mass[:, None] * np.abs(acc[:, :3])).sum()
    row = {"config": "c4", "n": int(n), "params": kw, "eps": args.eps,
           "scheme": args.scheme, "nimage": args.nimage if args.scheme == "images" else None,
           "rms_force": rms, "entries": {k: st["entries_" + k] for k in batch.KINDS},
           "pairs_by_kind": {k: st["pairs_" + k] for k in batch.KINDS},
           "host_lists_s": st["t_lists"], "groups": st["ngroups"], "pairs": st["pairs_total"],
           "kernel_ms_eager": eager, "ms_graph_replay_wall": graph,
           "interactions_per_s": st["pairs_total"] / (eager * 1e-3),
           "errors_vs_direct_ewald": {"median_over_rms": float(np.median(err) / rms),
                                      "p99_over_rms": float(np.percentile(err, 99) / rms),
                                      "median_rel": _errors(acc[sample], ref)["median"]},
           "net_force_over_sum_abs": float(net / typical)}
    _log("c4 %s: eager %.3f ms, graph %.3f ms, median |da|/rms %.2e, p99 %.2e" %
         (label, eager, graph, row["errors_vs_direct_ewald"]["median_over_rms"],
          row["errors_vs_direct_ewald"]["p99_over_rms"]))
    plan.close()
    print(json.dumps(row), flush=True)
    return ref, rms


def c4(args):
    pos, mass = ics.lattice_grf(args.ng, seed=4)
    n = pos.shape[0]
    sample = np.random.default_rng(7).choice(n, 1024, replace=False)
    dec = _decomposition(n, args.theta, "lattice")
    ref, rms = _c4_run(args, pos, mass, "theta %.2f" % args.theta, sample, None,
                       theta=args.theta, **dec)
    if args.mac == "tolerance":
        # errtol as a fraction of the rms force of the sampled bodies
        for f in args.errtols:
            _c4_run(args, pos, mass, "errtol %.0e rms" % f, sample, ref, mac="tolerance",
                    errtol=f * rms, merge_groups=dec["merge_groups"], leaf_max=dec["leaf_max"],
                    persistent_blocks=dec["persistent_blocks"])


def c5(args):
    pos, mass = ics.clustered(args.n, seed=31)
    n = pos.shape[0]
    rng = np.random.default_rng(5)
    w4 = rng.normal(size=(n, 4))
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4, soft="plummer", fp64=True)
    st = plan.stats()
    plan.evaluate("rel32c")
    fwd = plan.time("rel32c", reps=3)
    plan.adjoint(w4)  # warm-up
    t0 = time.perf_counter()
    grad = plan.adjoint(w4)
    adj = (time.perf_counter() - t0) * 1e3
    h = 1e-5
    picks = rng.choice(n, 16, replace=False)

    def loss():
        plan.evaluate("f64")
        return float((plan.result("f64") * w4).sum())

    def central(k, q, hh):
        p1, p2, m1, m2 = pos.copy(), pos.copy(), mass.copy(), mass.copy()
        if q < 3:
            p1[k, q] += hh
            p2[k, q] -= hh
            step = hh
        else:
            m1[k] *= 1 + hh
            m2[k] *= 1 - hh
            step = mass[k] * hh
        plan.update(p1, m1)
        lp = loss()
        plan.update(p2, m2)
        lm = loss()
        return (lp - lm) / (2 * step)

    # central differences at h and h/2; Richardson's combination removes the h^2
    # term, which at h = 1e-5 is not small next to bodies 1e-4 apart in the halos
    fd, fd2, ad = [], [], []
    for k in picks:
        for q in range(4):
            fd.append(central(k, q, h))
            fd2.append(central(k, q, h / 2))
            ad.append(grad[k, q])
    plan.update(pos, mass)
    fd, fd2, ad = (np.array(v).reshape(-1, 4) for v in (fd, fd2, ad))
    rich = (4 * fd2 - fd) / 3
    out = {}
    for q, name in ((slice(0, 3), "position"), (slice(3, 4), "mass")):
        scale = np.sqrt((rich[:, q] ** 2).mean())
        out[name] = float(np.abs(ad[:, q] - fd[:, q]).max() / scale)
        out[name + "_richardson"] = float(np.abs(ad[:, q] - rich[:, q]).max() / scale)
        out[name + "_richardson_median"] = float(np.median(np.abs(ad[:, q] - rich[:, q])) / scale)
    row = {"config": "c5", "n": int(n), "groups": st["ngroups"], "pairs": st["pairs_total"],
           "forward_kernel_ms": fwd, "adjoint_wall_ms": adj,
           "fd_max_err_over_rms": out, "fd_samples": int(picks.size)}
    _log("c5: forward %.3f ms, adjoint %.1f ms (wall, incl. transfers), FD %s" % (fwd, adj, out))
    plan.close()
    print(json.dumps(row), flush=True)


def dropin(args):
    import pgrav
    from pgrav import reflists
    pos, mass = ics.clustered(args.n, seed=3)
    n = pos.shape[0]
    eps = 1e-3
    kw = dict(theta=0.6, order=4, eps=eps, soft="k1", G=1.0)
    H = batch.HostPlan(pos, mass, **kw)
    perm, spos = H.sorted()
    st = H.stats()
    G = st["ngroups"]
    t0 = time.perf_counter()
    work = []
    for g in range(G):
        first, cnt, _ = H.group(g)
        soa, source_n = reflists.group_soa(H, g, "hexa")
        seg, nb = reflists.group_segments(H, g)
        work.append((first, cnt, soa, source_n, np.ascontiguousarray(seg.reshape(-1)), nb))
    _log("dropin: %d groups, lists in reference layouts %.1fs" % (G, time.perf_counter() - t0))
    btab = np.zeros((n, 14), np.float32)
    btab[:, 0] = mass[perm]
    btab[:, 1:4] = spos
    flat = btab.reshape(-1)
    gc = pgrav.GravCuda()
    times, results = {}, {}
    for mode in (0, 1):  # one launch per call, then the deferred batches
        pgrav.set_dropin_deferral(mode)
        for rep in range(3):
            btab[:, 7:11] = 0
            t1 = time.perf_counter()
            gc.walk_init_sink(flat, 14)
            for first, cnt, soa, source_n, seg, nb in work:
                q = -1
                while q < 0:
                    q, _ = gc.qalloc(count=False)
                if source_n:
                    gc.grav_mn("pH", first, cnt, soa, source_n, 29, 0.0, q)
                if nb:
                    gc.grav_mns("pMM_sK1", first, cnt, seg, nb, float(mass[0]), 1.0 / eps, q)
            gc.walk_terminate_sink(flat, 14)
            ms = (time.perf_counter() - t1) * 1e3
            times[mode] = min(times.get(mode, ms), ms)
            _log("dropin deferral %d rep %d: %.1f ms" % (mode, rep, ms))
        results[mode] = btab[:, 7:11].astype(np.float64)
    # the same loop with an empty library call in place of each call: the share of
    # the wall time that is Python and ctypes, not the library
    t1 = time.perf_counter()
    for first, cnt, soa, source_n, seg, nb in work:
        gc.qalloc(count=False)
        if source_n:
            gc.vecwidth()
        if nb:
            gc.vecwidth()
    loop_ms = (time.perf_counter() - t1) * 1e3
    best = times[1]
    got = results[1]
    plan = batch.Plan(pos, mass, fp64=True, **kw)
    plan.evaluate("abs32")
    kern = plan.time("abs32", reps=3)
    plan.evaluate("f64")
    ref = plan.result("f64")[perm]
    plan.close()
    da = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    pairs = st["pairs_hexa"] + st["pairs_body"]
    dd = np.linalg.norm(results[0][:, :3] - got[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
    row = {"config": "dropin", "n": int(n), "groups": G, "pairs_hexa": st["pairs_hexa"],
           "pairs_body": st["pairs_body"], "dropin_ms": best, "per_call_ms": times[0],
           "per_call_vs_deferred_max_rel": float(dd.max()), "python_loop_ms": loop_ms,
           "list_mb": float(sum(w[2].nbytes * (w[3] > 0) + 4 * w[5] for w in work) / 1e6),
           "dropin_interactions_per_s": pairs / (best * 1e-3), "calls": int(sum(
               (w[3] > 0) + (w[5] > 0) for w in work)),
           "batched_abs32_kernel_ms": kern, "batched_interactions_per_s": pairs / (kern * 1e-3),
           "median_rel_err_vs_fp64": float(np.median(da)), "max_rel_err_vs_fp64": float(da.max())}
    _log("dropin: %.1f ms (%.3g interactions/s) vs batched %.3f ms; median err %.2e" %
         (best, row["dropin_interactions_per_s"], kern, row["median_rel_err_vs_fp64"]))
    print(json.dumps(row), flush=True)


def dropin_sb(args):
    """The subtract_background walk's GPU work through the drop-in ABI: its pH / pQ
    calls, pQ1 flushes and pMM_mnss_sK1 flushes (the batched p-p queue), each class
    timed on its own from the first call to the end of WalkTerminateSinkCUDA."""
    import pgrav
    from pgrav.dlwalk import DLWalk
    pos, mass = ics.lattice_grf(args.ng, seed=4, rms=0.3)
    n = pos.shape[0]
    t0 = time.perf_counter()
    W = DLWalk(pos, mass, dl=dict(subtract_background=1, geometric_center=1, tail_spans=1),
               order=4, theta=args.theta, periodic="images", box=1.0, nimage=1)
    st, sb = W.stats(), W.sb_stats()
    _log("dropin_sb %d^3: walk %.1fs, %d pH + %d pQ calls, %d pQ1 flushes, %d pMM_mnss_sK1 "
         "flushes (%.3e pairs)" % (args.ng, time.perf_counter() - t0, st["ph_calls"],
                                   st["pq_calls"], st["pq1_flushes"], sb["mns_flushes"],
                                   sb["mns_pairs"]))
    body = W.bodies
    btab = np.zeros((n, 14), np.float32)
    btab[:, 0] = body[:, 0]
    btab[:, 1:4] = body[:, 1:4]
    flat = btab.reshape(-1)
    src = W.sources
    calls = [(("pH", "pQ")[r], int(f), int(m), np.ascontiguousarray(src[so:so + sn * sz]),
              int(sn), int(sz)) for r, f, m, so, sn, sz in W.calls]
    qfl = [(int(b), int(m), np.ascontiguousarray(W.qseg[s0:s0 + m]).reshape(-1),
            np.ascontiguousarray(src[so:so + sl * 13]), int(sl))
           for b, m, s0, so, sl, _ in W.qflush]
    mfl = [(int(b), int(m), W.mss_index[i0:i0 + m], W.mss_seg[e0:e0 + sl].reshape(-1),
            W.mseg[g0:g0 + sn].reshape(-1), W.mss_source_n[e0:e0 + sl])
           for b, m, i0, e0, sl, g0, sn, _ in W.mflush]
    gc = pgrav.GravCuda()

    def slot():
        q = -1
        while q < 0:
            q, _ = gc.qalloc(count=False)
        return q

    def run(kind):
        btab[:, 7:11] = 0
        t1 = time.perf_counter()
        gc.walk_init_sink(flat, 14)
        if kind in ("mn", "all"):
            for r, f, m, buf, sn, sz in calls:
                gc.grav_mn(r, f, m, buf, sn, sz, 0.0, slot())
        if kind in ("qnss", "all"):
            for b, m, seg, buf, sl in qfl:
                gc.grav_qnss("pQ1_qnss", b, m, seg, buf, sl, slot())
        if kind in ("mnss", "all"):
            for b, m, idx, ssg, seg, srcn in mfl:
                gc.grav_mnss("pMM_mnss_sK1", b, m, idx, ssg, seg, 0, srcn, float(body[0, 0]),
                             1e5, slot())
        gc.walk_terminate_sink(flat, 14)
        return (time.perf_counter() - t1) * 1e3
    row = {"config": "dropin_sb", "ng": args.ng, "n": int(n), "theta": args.theta,
           "ph_calls": st["ph_calls"], "pq_calls": st["pq_calls"], "pq1_flushes": st["pq1_flushes"],
           "mns_flushes": sb["mns_flushes"], "mns_pairs": sb["mns_pairs"],
           "mxn_h_pairs": st["mxn_h_pairs"], "mxn_q_pairs": st["mxn_q_pairs"], "ms": {}}
    for kind in ("all", "mn", "qnss", "mnss"):
        best = min(run(kind) for _ in range(3))
        row["ms"][kind] = best
        _log("   %-5s %9.2f ms" % (kind, best))
    row["mnss_pairs_per_s_wall"] = sb["mns_pairs"] / (row["ms"]["mnss"] * 1e-3)
    print(json.dumps(row), flush=True)


def steps(args):
    """The kernels either side of the force evaluation at the benchmark's N: the
    moment update under a fixed tree (pgrav_plan_update_device: body gather, one
    moment launch per level, record preparation) and the kick-drift of the
    reference's 56-byte body table (pgrav_integrate_device), each timed over 5
    calls with the inputs already on the device."""
    import torch
    from pgrav import integrate as integ
    pos, mass = ics.clustered(args.n, seed=2)
    n = pos.shape[0]
    dec = _decomposition(n, 0.6, "clustered")
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4, soft="plummer", fp64=False, **dec)
    st = plan.stats()
    rng = np.random.default_rng(7)
    moved = pos + 1e-6 * rng.standard_normal(pos.shape)
    dpos = torch.from_numpy(moved).cuda()
    dmass = torch.from_numpy(np.ascontiguousarray(mass, np.float64)).cuda()
    torch.cuda.synchronize()
    plan.update_device(dpos.data_ptr(), dmass.data_ptr())
    plan.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        plan.update_device(dpos.data_ptr(), dmass.data_ptr())
    plan.synchronize()
    upd = (time.perf_counter() - t0) / 5 * 1e3
    _log("steps: update (gather, per-level moments, records) %.2f ms per call, %d nodes" %
         (upd, st["nnodes"]))
    plan.close()
    tab = np.zeros(n, integ.BODY)
    tab["mass"] = mass
    tab["pos"] = pos
    tab["vel"] = 1e-3 * rng.standard_normal((n, 3))
    tab["acc"] = 1e-2 * rng.standard_normal((n, 3))
    dtab = torch.from_numpy(tab.view(np.uint8)).cuda()
    torch.cuda.synchronize()
    integ.integrate_device(dtab.data_ptr(), n, integ.BODY, 1e-3, 1e-3, 0.0, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        integ.integrate_device(dtab.data_ptr(), n, integ.BODY, 1e-3, 1e-3, 0.0, 0.0)
    torch.cuda.synchronize()
    itg = (time.perf_counter() - t0) / 5 * 1e3
    _log("steps: integrate %.3f ms per call (%d bodies x 56 B)" % (itg, n))
    print(json.dumps({"config": "steps", "n": int(n), "nodes": st["nnodes"], "update_ms": upd,
                      "integrate_ms": itg, "table_bytes": int(56 * n)}), flush=True)


def mac(args):
    pos, mass = ics.clustered(args.n, seed=2)
    rng = np.random.default_rng(12345)
    sample = rng.choice(pos.shape[0], min(8192, pos.shape[0]), replace=False)
    runs = [("theta", dict(theta=0.6, **_decomposition(pos.shape[0], 0.6, "clustered")))]
    ref = amed = None
    for et in args.errtols:
        for leaf in args.leaves:
            runs.append(("tolerance", dict(mac="tolerance", errtol=et, merge_groups=5, group_max=64,
                                           leaf_max=leaf, persistent_blocks=-1)))
    for kind, kw in runs:
        if kind == "tolerance":
            kw = dict(kw, errtol=kw["errtol"] * amed)
        t0 = time.perf_counter()
        plan = batch.Plan(pos, mass, order=4, eps=1e-4, soft="plummer", fp64=False, **kw)
        st = plan.stats()
        build = time.perf_counter() - t0
        if ref is None:
            ref = plan.direct(sample)
            amed = float(np.median(np.linalg.norm(ref[:, :3], axis=1)))
        plan.evaluate("rel32c")
        ms = plan.time("rel32c", reps=3)
        e = _errors(plan.result("rel32c")[sample], ref)
        plan.close()
        row = {"config": "mac", "n": int(pos.shape[0]), "mac": kind, "params": kw,
               "median_abs_acc": amed, "groups": st["ngroups"], "build_s": build,
               "pairs": {k: st["pairs_" + k] for k in ("hexa", "quad", "mono", "body")},
               "pairs_total": st["pairs_total"], "kernel_ms": ms,
               "interactions_per_s": st["pairs_total"] / (ms * 1e-3), "errors": e}
        _log("mac %-9s %s: %.3e pairs (h %.2e q %.2e m %.2e b %.2e) %8.2f ms  median %.2e p99 %.2e"
             % (kind, {k: v for k, v in kw.items() if k in ("theta", "errtol", "leaf_max")},
                st["pairs_total"], st["pairs_hexa"], st["pairs_quad"], st["pairs_mono"],
                st["pairs_body"], ms, e["median"], e["p99"]))
        print(json.dumps(row), flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("config", choices=["c2", "c3", "c4", "c5", "dropin", "dropin_sb", "mac",
                                       "steps"])
    ap.add_argument("--errtols", type=float, nargs="+", default=[1e-2, 1e-3, 1e-4])
    ap.add_argument("--leaves", type=int, nargs="+", default=[16, 64])
    ap.add_argument("--group-boxes", type=int, nargs="+", default=[0])
    ap.add_argument("--mac", choices=["theta", "tolerance"], default="theta")
    ap.add_argument("--ewald-ratio", type=float, default=0.05,
                    help="periodic 1: group radius / image distance below which a source's "
                         "Ewald correction joins the group's expansion (0: per pair)")
    ap.add_argument("--variants", nargs="+", default=["rel32c", "rel32", "abs32", "f64"])
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--offset", type=float, default=0.0,
                    help="c3: shift the system by this many system sizes")
    ap.add_argument("--ng", type=int, default=256)
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--thetas", type=float, nargs="+", default=None)
    ap.add_argument("--scheme", choices=["ewald", "images"], default="ewald")
    ap.add_argument("--nimage", type=int, default=2)
    ap.add_argument("--copy-check-mib", type=int, default=0,
                    help="first run a plain elementwise kernel over this many MiB (a known "
                         "byte count for calibrating the HBM traffic counters)")
    a = ap.parse_args(argv)
    a.thetas_given = a.thetas is not None
    a.thetas = a.thetas or [0.5, 0.6, 0.7]
    import torch  # noqa: F401  (one HIP runtime in the process: torch's)
    if a.copy_check_mib > 0:
        src = torch.ones(a.copy_check_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)  # reads and writes copy_check_mib MiB each
        torch.cuda.synchronize()
        _log("copy check: 3 x (read %d MiB, write %d MiB)" % (a.copy_check_mib, a.copy_check_mib))
        del src, dst
    if a.config == "c2":
        a.n = a.n or 2097152
        c2(a)
    elif a.config == "c3":
        a.n = a.n or 16777216
        c3(a)
    elif a.config == "c4":
        for theta in (a.thetas if a.thetas_given else [a.theta]):
            a.theta = theta
            c4(a)
    elif a.config == "dropin":
        a.n = a.n or 262144
        dropin(a)
    elif a.config == "dropin_sb":
        if a.ng == 256:
            a.ng = 32  # the host walk emits every body's leaf lists: keep it small
        dropin_sb(a)
    elif a.config == "mac":
        a.n = a.n or 16777216
        mac(a)
    elif a.config == "steps":
        a.n = a.n or 16777216
        steps(a)
    else:
        a.n = a.n or 1048576
        c5(a)


if __name__ == "__main__":
    main()
