"""Benchmark: batched mixed-precision gravity interaction-list evaluation on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--theta T] [--quick]

Workload (default): 16,777,216 clustered bodies (512 Plummer halos with
dN/dM ~ M^-1.9 over 3 decades, centres uniform in the unit cube, radii ~ M^1/3,
plus 20% uniform background), order-4 cells, theta 0.6, Plummer softening 1e-4,
G = 1.  One step = one force evaluation of every body's interaction list (one
kernel launch) with inputs resident in HBM.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "ahotsph-2hot-apps_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = ("interactions/s + % FP32 peak, mixed vs plain FP64 GPU at fixed force error, N=16.8M")
PEAK_FP32_TFLOPS = 157.3  # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0
# reference-stated flops per interaction (nln/do_grav_sse4.c:45, 388, 1046)
FLOPS = {"hexa": 217, "quad": 66, "mono": 26, "body": 26}
# algorithmic bytes per list entry / per sink (prepared FP32 record + FP64 centre)
BYTES_ENTRY = {"hexa": 4 + 32 + 184, "quad": 4 + 32 + 40, "mono": 4 + 32, "body": 4 + 32}
BYTES_SINK = 32 + 16


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(args):
    from pgrav import ics
    if args.workload == "clustered":
        return ics.clustered(args.n, seed=args.seed)
    if args.workload == "plummer":
        return ics.plummer(args.n, seed=args.seed)
    ng = int(round(args.n ** (1 / 3)))
    return ics.lattice_grf(ng, seed=args.seed)


def workload_name(args, n, variant=None, dec=None):
    name = "%s N=%d order %d theta %.2f Plummer eps %g G 1" % (args.workload, n, args.order,
                                                              args.theta, args.eps)
    if variant is not None:  # key of a measured-traffic record: the launch too
        name += " | %s groups %d/%d/%d persistent %d" % (
            variant, dec["merge_groups"], dec["group_max"], dec["leaf_max"],
            dec["persistent_blocks"])
        if dec.get("body_unroll", 4) != 4:
            name += " unroll %d" % dec["body_unroll"]
        if dec.get("tile", 64) != 64:
            name += " tile %d" % dec["tile"]
        if dec.get("dispatch", 2) != 2:
            name += " dispatch %d" % dec["dispatch"]
    return name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16777216)
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--workload", default="clustered", choices=["clustered", "plummer", "lattice"])
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--variant", default="rel32c")
    ap.add_argument("--group-mode", type=int, default=1)
    ap.add_argument("--leaf", type=int, default=16, help="leaf size without tuned.json")
    ap.add_argument("--persistent", type=int, default=-1)
    ap.add_argument("--sample", type=int, default=8192)
    ap.add_argument("--no-plain64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tuned", action="store_true", help="ignore tuned.json")
    ap.add_argument("--quick", action="store_true", help="N = 2^21 for a fast check")
    ap.add_argument("--copy-check-mib", type=int, default=0,
                    help="first run an elementwise kernel that reads and writes this many "
                         "MiB, three times (the known-size kernel that scales rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: there is no multi-GPU path; rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.quick:
        args.n = 1 << 21

    import torch
    assert torch.cuda.is_available(), "no GPU"
    torch.cuda.init()
    from pgrav import batch

    if args.copy_check_mib > 0:
        n4 = args.copy_check_mib * (1 << 20) // 4
        src = torch.ones(n4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 2.0, out=dst)  # reads and writes copy_check_mib MiB each
        torch.cuda.synchronize()
        log("copy check: 3 x (read %d MiB, write %d MiB)" % (args.copy_check_mib,
                                                           args.copy_check_mib))
        del src, dst

    t0 = time.perf_counter()
    pos, mass = make_workload(args)
    log("workload %s N=%d built in %.1fs" % (args.workload, pos.shape[0], time.perf_counter() - t0))
    # decomposition parameters: the autotuner's choice for this workload when
    # ahotsph-2hot-apps_amd/tuned.json holds one
    from pgrav import autotune
    key = autotune.workload_key(pos.shape[0], args.theta, args.order, "plummer", args.eps,
                                args.workload)
    tuned = None if args.no_tuned else autotune.load(key)
    dec = {"merge_groups": args.group_mode, "group_max": 64, "leaf_max": args.leaf,
           "persistent_blocks": args.persistent, "body_unroll": 4, "tile": 64, "dispatch": 2}
    if tuned:
        dec.update({k: tuned["best"][k] for k in dec if k in tuned["best"]})
        log("using tuned decomposition %s" % dec)
    t0 = time.perf_counter()
    plan = batch.Plan(pos, mass, theta=args.theta, order=args.order, eps=args.eps,
                      soft="plummer", merge_groups=dec["merge_groups"],
                      group_max=dec["group_max"], leaf_max=dec["leaf_max"],
                      persistent_blocks=dec["persistent_blocks"], body_unroll=dec["body_unroll"],
                      dispatch=dec["dispatch"], fp64=True)
    if dec["tile"] != 64:
        plan.set_tile(dec["tile"])
    st = plan.stats()
    log("plan: %.1fs (tree %.1fs lists %.1fs upload %.1fs) groups %d mean %.1f pairs %.3e" %
        (time.perf_counter() - t0, st["t_tree"], st["t_lists"], st["t_upload"], st["ngroups"],
         st["mean_group"], st["pairs_total"]))

    var = args.variant
    # a step replays the evaluation captured once in a hipGraph (the persistent
    # launch resets its own work counters, so replays need no host work)
    for _ in range(args.warmup):
        plan.evaluate(var)
    plan.graph_capture(var)
    plan.graph_launch()
    plan.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.graph_launch()
    plan.synchronize()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / args.steps
    # device time of the kernel itself: HIP events around back-to-back launches
    # on the stream the kernel runs on
    kern_ms = plan.time(var, reps=max(1, args.steps))
    log("step %.3f ms (wall, graph replay)  kernel %.3f ms (HIP events)" % (wall * 1e3, kern_ms))

    pairs = {k: st["pairs_" + k] for k in FLOPS}
    npairs = sum(pairs.values())
    flops = sum(FLOPS[k] * pairs[k] for k in FLOPS)
    nbytes = sum(BYTES_ENTRY[k] * st["entries_" + k] for k in FLOPS) + BYTES_SINK * st["n"]
    achieved = flops / (kern_ms * 1e-3) / 1e12

    # accuracy at this configuration vs the FP64 direct sum on sampled bodies
    rng = np.random.default_rng(12345)
    sample = rng.choice(pos.shape[0], min(args.sample, pos.shape[0]), replace=False)
    ref = plan.direct(sample)
    acc = plan.result(var)
    da, dp = batch.rel_errors(acc[sample], ref)
    errors = {"median": float(np.median(da)), "p90": float(np.percentile(da, 90)),
              "p99": float(np.percentile(da, 99)), "phi_median": float(np.median(dp)),
              "sample": int(sample.size)}
    log("errors |da|/|a|: median %.2e p90 %.2e p99 %.2e" % (errors["median"], errors["p90"],
                                                           errors["p99"]))

    baselines = {}
    f64_ms = plan.time("f64", reps=1)
    baselines["batched_fp64_ms"] = f64_ms
    if not args.no_plain64:
        plain_ms = plan.time("plain64", reps=1)
        baselines["plain_fp64_one_launch_per_group_ms"] = plain_ms
        baselines["speedup_vs_plain_fp64"] = plain_ms / kern_ms
        log("plain FP64 one-launch-per-group %.1f ms -> speedup %.1fx" % (plain_ms, plain_ms / kern_ms))
    baselines["speedup_vs_batched_fp64"] = f64_ms / kern_ms

    # HBM bytes per launch of the batched kernel, measured with rocprofv3
    # FETCH_SIZE / WRITE_SIZE for this exact workload and scaled by the same
    # counters on a copy of known size (profiles/traffic.json records how)
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(workload_name(args, pos.shape[0], var, dec))
        if rec:
            traffic = rec["bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        traffic = None

    # FP32 flops the hardware issued per launch for this exact workload and launch
    # (rocprofv3 SQ_INSTS_VALU_FLOPS_FP32 x 64 lanes; profiles/valu_flops.json
    # records the run): the frac by the hardware's count beside the reference's
    hw_flops = None
    try:
        with open(os.path.join(ROOT, "profiles", "valu_flops.json")) as f:
            rec = json.load(f).get(workload_name(args, pos.shape[0], var, dec))
        if rec:
            hw_flops = rec["fp32_flops_per_launch"]
    except (OSError, ValueError, KeyError):
        hw_flops = None

    cpu = None
    if not args.no_cpu:
        # the reference's own SSE4/AVX CPU kernels (oracle/_ref) on a bounded
        # sample of the same workload's groups; the checker side, not the product
        try:
            from oracle.cpu_baseline import run_cpu_baseline
            cpu = run_cpu_baseline(pos, mass, args, acc_gpu=acc, dec=dec)
            log("cpu baseline: %.3e interactions/s on %d threads (%s)" %
                (cpu["value"], cpu["cores"], cpu["cpu_model"]))
        except Exception as e:  # reported, never silently replaced
            log("cpu baseline unavailable: %s" % e)
            cpu = {"error": str(e)}

    out = {
        "metric": METRIC,
        "value": npairs / (kern_ms * 1e-3),
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": workload_name(args, pos.shape[0]),
                   "variant": var, "ngroups": st["ngroups"], "pairs": pairs,
                   "mean_group": st["mean_group"], "decomposition": dec,
                   "host_build_s": st["t_tree"] + st["t_lists"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel_ms": kern_ms, "algorithmic_bytes": nbytes,
                     "hbm_GBs_algorithmic": nbytes / (kern_ms * 1e-3) / 1e9,
                     "hw_fp32_flops_per_launch": hw_flops,
                     "hw_frac": (hw_flops / (kern_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS
                                 if hw_flops else None)},
        "pct_fp32_peak": 100.0 * achieved / PEAK_FP32_TFLOPS,
        "errors": errors,
        "baselines": baselines,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
