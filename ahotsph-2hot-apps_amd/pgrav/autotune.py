"""Autotuner for the batched evaluation.

Searches the parameters that change the work decomposition but not the
physics: how sink groups are formed (tree cells with siblings merged, then
optionally merged further where the joint list costs less, or cells cut into
balanced groups by coordinate bisection), the group size
(lanes per wave used), the leaf size, the launch shape (one workgroup per group in
longest-first order, or a persistent grid of k single-wave workgroups per CU
draining a work counter), the group order inside the per-XCD slices, and the
unroll factor of the kernel's order-0 and body loops and the tile of list
entries staged in LDS per round (compiled variants).  The opening angle, multipole order and softening are inputs, so the
accuracy target is held fixed.  Each candidate is timed with HIP events over
back-to-back launches on the plan's stream; the fastest wins.

    python -m pgrav.autotune --n 2097152 --out tuned.json

Results are stored per workload key in a JSON file that bench.py reads.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

from . import batch

DEFAULT_SPACE = {
    "merge_groups": [1, 4, 5],
    "group_max": [48, 64],
    "leaf_max": [16, 32, 48, 64],
}
# (unroll, tile) of the rel32c kernels (compiled variants, chosen per plan without
# rebuilding it): the unroll of the order-0 / body loops, and the list entries
# staged in LDS per round (one barrier pair and one compensated fold per round);
# tiles of 32 and 128 are compiled with the unroll of 4
UNROLL_TILE = [(2, 64), (4, 64), (8, 64), (4, 32), (4, 128)]
# group order inside the per-XCD slices: 1 longest-first, 2 Morton order (a
# plan-level choice, no rebuild)
DISPATCH = [1, 2]
# launch shapes: 0 = one workgroup per group; -1 = persistent grid of 32
# single-wave workgroups per CU; k > 0 = persistent grid of k per CU.  The
# table and tuned.json store the resulting workgroup count (what
# pgrav_plan_set_launch and Plan(persistent_blocks=...) take).
PERSISTENT = [0, -1]
TUNED = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuned.json")


def workload_key(n, theta, order, soft, eps, workload):
    return "%s:n=%d:theta=%.3g:order=%d:soft=%s:eps=%.3g" % (workload, n, theta, order, soft, eps)


def tune(pos, mass, base, space=None, reps=3, variant="rel32c", log=print):
    """Returns (best_params, table).  `base` holds theta/order/eps/soft/G/..."""
    space = space or DEFAULT_SPACE
    keys = sorted(space)
    table = []
    best = None
    for combo in itertools.product(*(space[k] for k in keys)):
        kw = dict(base)
        kw.update(dict(zip(keys, combo)))
        t0 = time.perf_counter()
        plan = batch.Plan(pos, mass, fp64=False, **kw)
        build = time.perf_counter() - t0
        st = plan.stats()
        plan.evaluate(variant)  # warm-up
        ut = UNROLL_TILE if variant == "rel32c" and not kw.get("periodic") in (1, True, "ewald") \
            else [(4, 64)]
        for pb, (bu, tile), dsp in itertools.product(PERSISTENT, ut, DISPATCH):
            blocks = pb * _num_cus() if pb > 0 else pb
            plan.set_unroll(bu)
            plan.set_tile(tile)
            plan.set_dispatch(dsp)
            ms = _time_with_blocks(plan, variant, reps, blocks)
            row = dict(kw, persistent_blocks=blocks, body_unroll=bu, tile=tile, dispatch=dsp, ms=ms,
                       pairs=st["pairs_total"], ngroups=st["ngroups"],
                       mean_group=st["mean_group"], build_s=build)
            table.append(row)
            log("tune %s -> %.3f ms (groups %d, mean %.1f, pairs %.3e)" %
                ({k: row[k] for k in keys + ["persistent_blocks", "body_unroll", "tile",
                                             "dispatch"]}, ms,
                 st["ngroups"], st["mean_group"], st["pairs_total"]))
            if best is None or ms < best["ms"]:
                best = row
        plan.close()
    return best, table


def _num_cus():
    import torch
    return torch.cuda.get_device_properties(0).multi_processor_count


def _time_with_blocks(plan, variant, reps, pb):
    from ._lib import check, lib
    import ctypes
    lib().pgrav_plan_set_launch(plan.h, pb)
    check()
    ms = ctypes.c_double(0)
    lib().pgrav_plan_time(plan.h, batch._v(variant), reps, None, ctypes.byref(ms))
    check()
    return ms.value


def load(key, path=TUNED):
    try:
        with open(path) as f:
            return json.load(f).get(key)
    except (OSError, ValueError):
        return None


def save(key, best, table, path=TUNED, merge=False):
    """Stores the best row and the table under `key`; merge: adds the rows to the
    stored table and keeps the faster of the stored and the new best."""
    data = {}
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    keep = ("merge_groups", "group_max", "leaf_max", "persistent_blocks", "body_unroll", "tile",
            "dispatch", "ms")
    rows = [{k: r[k] for k in keep + ("pairs", "mean_group") if k in r} for r in table]
    best = {k: best[k] for k in keep if k in best}
    if merge and key in data:
        rows = data[key].get("table", []) + rows
        if data[key]["best"]["ms"] < best["ms"]:
            best = data[key]["best"]
    data[key] = {"best": best, "table": rows}
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2097152)
    ap.add_argument("--workload", default="clustered")
    ap.add_argument("--theta", type=float, default=0.6)
    ap.add_argument("--order", type=int, default=4)
    ap.add_argument("--eps", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--out", default=TUNED)
    ap.add_argument("--refine", action="store_true",
                    help="keep the stored best decomposition (groups, leaves) and search only "
                         "the launch choices (launch shape, unroll, tile, dispatch); rows are "
                         "added to the stored table")
    a = ap.parse_args(argv)
    import torch  # noqa: F401  (one HIP runtime in the process: torch's)
    from . import ics
    if a.workload == "clustered":
        pos, mass = ics.clustered(a.n, seed=a.seed)
    elif a.workload == "plummer":
        pos, mass = ics.plummer(a.n, seed=a.seed)
    else:
        pos, mass = ics.lattice_grf(int(round(a.n ** (1 / 3))), seed=a.seed)
    base = dict(theta=a.theta, order=a.order, eps=a.eps, soft="plummer",
                periodic=a.workload == "lattice")
    key = workload_key(pos.shape[0], a.theta, a.order, "plummer", a.eps, a.workload)
    space = None
    if a.refine:
        stored = load(key, a.out)
        if not stored:
            raise SystemExit("--refine: no stored decomposition for " + key)
        space = {k: [stored["best"][k]] for k in DEFAULT_SPACE}
    best, table = tune(pos, mass, base, space=space,
                       log=lambda s: print(s, file=sys.stderr, flush=True))
    save(key, best, table, a.out, merge=a.refine)
    best = load(key, a.out)["best"]
    print(json.dumps({"key": key, "best": best}))


if __name__ == "__main__":
    main()
