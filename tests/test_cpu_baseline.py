"""The CPU baseline harness (oracle/cpu_baseline.*) runs the reference's own
SSE4/AVX kernels over the plan's lists; its per-sink results must match the
FP64 list oracle on the same groups, or the rate it reports is not a rate of
the same work."""
import numpy as np
import pytest

import oracle
from oracle import cpu_baseline
from pgrav import batch, ics
from listoracle import group_reference

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="oracle/_ref not built (needs the reference sources)")


@pytest.mark.parametrize("mode", [1, 5])
@pytest.mark.parametrize("order,kind,cell,csz", [(4, "hexa", "ph", 29), (2, "quad", "pq", 13)])
@pytest.mark.parametrize("arch", ["sse4", "avx8"])
def test_harness_results_match_list_oracle(arch, order, kind, cell, csz, mode):
    pos, mass = ics.clustered(6000, seed=5)
    H = batch.HostPlan(pos, mass, theta=0.6, order=order, eps=1e-3, soft="plummer",
                       merge_groups=mode, rcb_cell=1024)
    st = H.stats()
    perm, spos = H.sorted()
    smass = mass[perm]
    # the harness takes every group as a run of sinks: the sinks go in group order
    # (bisection groups, mode 5, are body sets of the sorted order)
    goff, gslots = H.group_slot_table()
    assert np.array_equal(np.sort(gslots), np.arange(st["n"]))
    first, count = goff[:-1].copy(), np.diff(goff).astype(np.int32)
    if mode == 1:
        assert np.array_equal(first, H.groups()[0])
    hoff, hidx = H.csr(kind)
    boff, bidx = H.csr("body")
    assert hoff[-1] == st["entries_" + kind] and boff[-1] == st["entries_body"]
    rec = H.records(np.arange(st["nnodes"], dtype=np.uint32))[:, :csz]
    D = {"rec": np.ascontiguousarray(rec, np.float32),
         "sinks": np.ascontiguousarray(spos[gslots], np.float32),
         "bodies": np.ascontiguousarray(np.concatenate([smass[:, None], spos], 1), np.float32),
         "first": first, "count": count, "hoff": hoff, "hidx": hidx, "boff": boff, "bidx": bidx}
    groups = np.arange(0, first.size, 7, dtype=np.int32)
    pin = cpu_baseline.host_cpu()["pin_cpus"]
    wall, ks, out = cpu_baseline._run_lists(arch, D, groups, 1e-6, 4, cell, csz,
                                            pin if len(pin) >= 4 else None)
    assert wall > 0 and ks[0] > 0 and ks[1] > 0  # cell and body kernel seconds
    # do_gravpq's quadrupole potential carries the sign pgrav.cu has (the -1
    # convention); do_gravph's is the physical one
    qsign = -1 if order == 2 else 1
    for g in groups[:12]:
        _, ref = group_reference(H, int(g), spos, smass, soft="plummer", eps=1e-3,
                                 quad_phi_sign=qsign)
        got = out[first[g]:first[g] + count[g]]
        rel = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1) / np.linalg.norm(ref[:, :3], axis=1)
        # FP32 absolute coordinates in the reference kernels: ~1e-5 relative
        assert np.median(rel) < 1e-4, (g, np.median(rel))
        assert np.allclose(got[:, 3], ref[:, 3], rtol=1e-4)
    # groups not in the sample are left untouched
    mask = np.ones(out.shape[0], bool)
    for g in groups:
        mask[first[g]:first[g] + count[g]] = False
    assert not out[mask].any()


def test_host_cpu_topology():
    """One pinned CPU per physical core of the affinity set, within the quota."""
    c = cpu_baseline.host_cpu()
    assert 1 <= len(c["pin_cpus"]) <= c["affinity_cpus"] <= c["nproc"]
    assert len(set(c["pin_cpus"])) == len(c["pin_cpus"])
    assert c["physical_cores"] >= len(c["pin_cpus"])
    if c["cgroup_cpu_quota"]:
        assert len(c["pin_cpus"]) <= c["cgroup_cpu_quota"]
    assert c["cpu_model"]
