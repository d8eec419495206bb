"""Parity of the batched path (pgrav_plan_*) on the GPU.

1. The exact lists of sampled groups are re-evaluated by the FP64 oracle (cell
   records rebuilt from each cell's bodies); the kernel variants must agree:
   FP64 to 1e-10, FP32 group-relative to 3e-6 (relative per sink, |da|/|a|).
2. The library's FP64 direct sum equals the oracle's FP64 direct sum (1e-12).
3. Graph replays, the persistent grid and the one-launch-per-group FP64
   dispatch reproduce the eager batched result.
"""
import numpy as np
import pytest

import listoracle
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu_lib):
    from pgrav import batch, ics
    pos, mass = ics.clustered(40000, seed=21, nhalo=24)
    return batch, ics, pos, mass


def _group_errors(batch, pos, mass, plan_kw, variants, ngroups=12, seed=0):
    plan = batch.Plan(pos, mass, **plan_kw)
    H = batch.HostPlan(pos, mass, **plan_kw)
    perm, spos = H.sorted()
    smass = mass[perm]
    out = {}
    for v in variants:
        plan.evaluate(v)
        out[v] = plan.result(v)[perm]  # back to sorted order
    st = H.stats()
    rng = np.random.default_rng(seed)
    errs = {v: [] for v in variants}
    for g in rng.choice(st["ngroups"], ngroups, replace=False):
        first, ref = listoracle.group_reference(H, int(g), spos, smass,
                                                soft=plan_kw.get("soft", "plummer"),
                                                eps=plan_kw.get("eps", 1e-4),
                                                G=plan_kw.get("G", 1.0))
        slots = H.group_slots(int(g))  # a run of bodies, or a bisection group's body set
        for v in variants:
            got = out[v][slots]
            da = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
            dp = np.abs(got[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])
            errs[v].append(max(da.max(), dp.max()))
    plan.close()
    return {v: max(e) for v, e in errs.items()}


@pytest.mark.parametrize("soft,eps,mode", [("plummer", 1e-3, 1), ("k1", 2e-3, 1), ("none", 0.0, 1),
                                           ("plummer", 1e-3, 5), ("k1", 2e-3, 5)])
def test_batched_variants_match_oracle_on_same_lists(setup, soft, eps, mode):
    """Every variant against the oracle's FP64 evaluation of the group's own
    lists; mode 5: sink groups from coordinate bisection (body sets, read and
    written through the slot table)."""
    batch, ics, pos, mass = setup
    kw = dict(theta=0.6, order=4, eps=eps, soft=soft, G=1.0, merge_groups=mode, leaf_max=32,
              rcb_cell=4096)
    e = _group_errors(batch, pos, mass, kw, ["rel32c", "rel32", "f64", "abs32"])
    assert e["f64"] < 1e-10, e
    # FP32 sums of ~4000 terms per sink: a few 1e-6 relative
    assert e["rel32c"] < 1e-5, e
    assert e["rel32"] < 1e-5, e
    assert e["abs32"] < 1e-3, e  # absolute FP32 coordinates: the precision loss the
    #                             group-relative path removes


@pytest.mark.parametrize("mode", [1, 5])
def test_tolerance_mac_lists_match_oracle(setup, mode):
    """Lists of the error-tolerance acceptance (cells at orders 0, 2 and 4 mixed
    by their critical radii, nln/mac.c:1776-1800): every variant against the
    oracle's FP64 evaluation of the same lists."""
    batch, ics, pos, mass = setup
    kw = dict(order=4, eps=1e-3, soft="plummer", G=1.0, merge_groups=mode, leaf_max=32,
              rcb_cell=4096, mac="tolerance", errtol=0.1, frac_tol=0.02)
    st = batch.HostPlan(pos, mass, **kw).stats()
    assert min(st["pairs_hexa"], st["pairs_quad"], st["pairs_mono"]) > 0, st
    e = _group_errors(batch, pos, mass, kw, ["rel32c", "f64"])
    assert e["f64"] < 1e-10 and e["rel32c"] < 1e-5, e


def test_tolerance_mac_error_bound_on_gpu(setup):
    """Each sink's error against the FP64 direct sum stays below (cells listed) x
    errtol: every accepted cell's expansion error is within errtol."""
    batch, ics, pos, mass = setup
    errtol = 0.05
    kw = dict(order=4, eps=1e-3, soft="plummer", leaf_max=32, mac="tolerance", errtol=errtol,
              merge_groups=5, rcb_cell=4096)
    plan = batch.Plan(pos, mass, **kw)
    H = batch.HostPlan(pos, mass, **kw)
    plan.evaluate("rel32c")
    got = plan.result("rel32c")
    sample = np.random.default_rng(8).choice(pos.shape[0], 512, replace=False)
    ref = plan.direct(sample)
    plan.close()
    # cells listed for each sampled body's group
    perm, _ = H.sorted()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size)
    off, slots = H.group_slot_table()
    group_of = np.empty(perm.size, np.int64)
    group_of[slots] = np.repeat(np.arange(off.size - 1), np.diff(off))
    ncell = sum(np.diff(H.csr(k)[0]) for k in ("hexa", "quad", "mono"))
    n_s = ncell[group_of[inv[sample]]]
    err = np.linalg.norm(got[sample, :3] - ref[:, :3], axis=1)
    # FP32 rounding adds ~1e-6 of |a| on top of the truncation bound
    assert np.all(err <= n_s * errtol + 1e-5 * np.linalg.norm(ref[:, :3], axis=1)), \
        (err / (n_s * errtol)).max()


@pytest.mark.parametrize("order", [2, 0])
def test_lower_orders_match_oracle(setup, order):
    batch, ics, pos, mass = setup
    kw = dict(theta=0.5, order=order, eps=1e-3, soft="plummer", G=2.0)
    e = _group_errors(batch, pos, mass, kw, ["rel32c", "f64"], ngroups=8)
    assert e["f64"] < 1e-10 and e["rel32c"] < 1e-5, e


def test_direct_sum_matches_oracle(setup):
    batch, ics, pos, mass = setup
    for soft, st, eps in (("plummer", 1, 1e-3), ("k1", 4, 3e-3), ("none", 0, 0.0)):
        plan = batch.Plan(pos, mass, eps=eps, soft=soft, G=1.5)
        sample = np.arange(0, pos.shape[0], 997)
        got = plan.direct(sample)
        ref = 1.5 * oracle.direct(pos[sample], pos, mass, soft_type=st, eps=eps)
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=0)
        plan.close()


def test_accuracy_against_direct_sum(setup):
    """Order-4 cells at theta = 0.5: the force error is the expansion error."""
    batch, ics, pos, mass = setup
    plan = batch.Plan(pos, mass, theta=0.5, order=4, eps=1e-4)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    sample = np.random.default_rng(3).choice(pos.shape[0], 1024, replace=False)
    da, dp = batch.rel_errors(acc[sample], plan.direct(sample))
    assert np.median(da) < 1e-4, np.median(da)
    assert np.percentile(da, 99) < 2e-3
    plan.close()


def test_graph_persistent_and_plain_dispatch_reproduce_eager(setup):
    batch, ics, pos, mass = setup
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4)
    plan.evaluate("rel32c")
    eager = plan.result("rel32c")
    plan.graph_capture("rel32c")
    for rep in range(3):
        plan.evaluate("f64")  # unrelated eager work between replays
        plan.synchronize()
        plan.graph_launch()
        np.testing.assert_array_equal(plan.result("rel32c"), eager)
    plan.close()
    pers = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4, persistent_blocks=-1)
    for rep in range(2):
        pers.evaluate("rel32c")
        np.testing.assert_array_equal(pers.result("rel32c"), eager)
    pers.graph_capture("rel32c")
    for rep in range(3):
        pers.graph_launch()
        np.testing.assert_array_equal(pers.result("rel32c"), eager)
    pers.evaluate("f64")
    f64 = pers.result("f64")
    pers.evaluate("plain64")
    plain = pers.result("plain64")
    np.testing.assert_allclose(plain, f64, rtol=1e-12, atol=1e-12 * np.abs(f64).max())
    pers.close()


@pytest.mark.parametrize("mode", [1, 5])
def test_unroll_variants_give_identical_results(setup, mode):
    """The autotuner's unroll choices (order-0 and body loops of rel32c) and the
    dispatch orders are the same arithmetic in the same order per group:
    bit-identical results, eager and from a graph, persistent (with its work
    stealing across XCD slices) and one workgroup per group."""
    batch, ics, pos, mass = setup
    plan = batch.Plan(pos, mass, theta=0.6, order=4, eps=1e-4, merge_groups=mode, rcb_cell=4096,
                      mac="tolerance", errtol=0.1)  # lists with order-0, 2, 4 cells and bodies
    st = plan.stats()
    assert min(st["pairs_mono"], st["pairs_body"]) > 0
    plan.evaluate("rel32c")
    want = plan.result("rel32c")
    for bu in (2, 8, 4):
        plan.set_unroll(bu)
        for pb in (0, -1):
            plan.set_launch(pb)
            plan.evaluate("rel32c")
            np.testing.assert_array_equal(plan.result("rel32c"), want)
    # the group orders (global longest-first; XCD slices ordered longest-first,
    # Morton, costliest-first then Morton) change only who evaluates a group when
    for mode in (0, 1, 3, 2):
        plan.set_dispatch(mode)
        plan.evaluate("rel32c")
        np.testing.assert_array_equal(plan.result("rel32c"), want)
    plan.set_unroll(8)
    plan.graph_capture("rel32c")
    for rep in range(2):
        plan.evaluate("f64")
        plan.synchronize()
        plan.graph_launch()
        np.testing.assert_array_equal(plan.result("rel32c"), want)
    from pgrav._lib import PgravError
    with pytest.raises(PgravError):
        plan.set_unroll(3)
    plan.close()


@pytest.mark.parametrize("mode,periodic", [(1, 0), (5, 0), (1, 2)])
def test_tile_variants(setup, mode, periodic):
    """The autotuner's tile (list entries per LDS round: 32, 64, 128) changes only
    where the compensated tile sums are folded: every tile matches the FP64
    evaluation of the same lists to FP32 accuracy, repeats bit for bit, eager and
    from a graph, and the periodic-1 kernels keep their 64."""
    batch, ics, pos, mass = setup
    kw = dict(order=4, eps=1e-3, merge_groups=mode, leaf_max=32, rcb_cell=4096, mac="tolerance",
              errtol=0.1, frac_tol=0.02)
    if periodic:
        p = (pos - pos.min(0)) / (np.ptp(pos, 0).max() * 1.001)
        kw.update(periodic="images", nimage=1, box=1.0)
    else:
        p = pos
    plan = batch.Plan(p, mass, **kw)
    plan.evaluate("f64")
    ref = plan.result("f64")
    out, err = {}, {}
    for tile in (32, 128, 64):
        plan.set_tile(tile)
        for pb in (0, -1):
            plan.set_launch(pb)
            plan.evaluate("rel32c")
            r = plan.result("rel32c")
            if tile in out:
                np.testing.assert_array_equal(r, out[tile])
            out[tile] = r
        da, _ = batch.rel_errors(out[tile], ref)
        dp = np.abs(out[tile][:, 3] - ref[:, 3]) / np.median(np.abs(ref[:, 3]))
        err[tile] = (np.percentile(da, 99.9), da.max(), dp.max())
    print("tile errors against f64 (99.9th pct and max |da|/|a|, max |dphi| / median |phi|):",
          err, flush=True)
    # every body (not sampled groups): the largest relative error sits at a sink
    # whose pull nearly cancels, the same body and size for every tile (the periodic
    # potential crosses zero, hence its scale)
    for tile, e in err.items():
        assert e[0] < 1e-5 and e[1] < 1e-4 and e[2] < 1e-5, (tile, err)
        assert e[1] < 1.5 * err[64][1] + 1e-6, (tile, err)
    plan.set_tile(128)
    plan.graph_capture("rel32c")
    for rep in range(2):
        plan.evaluate("f64")
        plan.synchronize()
        plan.graph_launch()
        np.testing.assert_array_equal(plan.result("rel32c"), out[128])
    from pgrav._lib import PgravError
    with pytest.raises(PgravError):
        plan.set_tile(96)
    plan.close()


def test_plan_rejects_bad_input(gpu_lib):
    from pgrav import batch
    from pgrav._lib import PgravError
    pos = np.random.default_rng(0).uniform(size=(100, 3))
    with pytest.raises(PgravError):
        batch.Plan(pos, np.full(100, 0.01), soft="plummer", eps=0.0)
    with pytest.raises(PgravError):
        batch.Plan(pos, np.full(100, 0.01), soft=3)
