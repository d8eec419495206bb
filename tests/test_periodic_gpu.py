"""Periodic evaluation (nearest images + the FP64 Ewald table) on the GPU.

1. Same lists: sampled groups' own lists re-evaluated by the oracle
   (listoracle.group_reference_periodic: cells rebuilt from their bodies, the
   Ewald correction from oracle.ewald_corr in the tabulated form the kernels
   use -- nearest node and its Taylor steps -- built from the exact lattice
   sum of nln/ewald.c:13-64 and its closed-form derivatives), both ways the
   kernels take it: per pair, and through the group's second-order expansion
   about its centre.  The FP64 kernel agrees to 1e-9 of the rms force and
   potential, the FP32 kernel to FP32 accuracy.  Planted mistakes in the oracle
   -- the correction scaled by 1 + 1e-3, an off-diagonal Jacobian sign flipped
   -- sit far outside the tolerance.
2. End to end: accelerations and potentials against the oracle's FP64 Ewald
   direct sum (nln/grav_n2.c pattern).  The reference's ewald() gives a body's
   own pair the constant pi / (alpha^2 L^3); the kernels give it the limit of
   the correction at r = 0, 2 x 1.4186487533 / L (its own images and the
   background): the potentials differ by m_i times the difference.
"""
import numpy as np
import pytest

import listoracle
import oracle

pytestmark = pytest.mark.gpu

XI = 2.837297479  # lim_{r->0} phi_ewald(r) + 1/r, unit box (2 x 1.4186487533, nln/ewald_le2.c:38)
TOL = 1e-9  # FP64 kernels against the oracle on the same lists, relative to the rms


@pytest.mark.parametrize("ratio", [0.0, 0.05])
def test_perfect_lattice_has_no_force(gpu_lib, ratio):
    from pgrav import batch
    ng = 12
    g = (np.arange(ng) + 0.5) / ng
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([gx, gy, gz], -1).reshape(-1, 3)
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.5, order=4, soft="none", eps=0.0,
                      ewald_ratio=ratio)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    pair = mass[0] * ng ** 2  # one neighbour's pull at one lattice spacing
    # the exact answer is 0; what remains is the theta = 0.5 expansion error of the
    # cells (an FP64 CPU evaluation of the same lists with exact ewald() per pair
    # leaves 1.4% of a pair force)
    assert np.abs(acc[:, :3]).max() < 3e-2 * pair, np.abs(acc[:, :3]).max() / pair
    plan.close()


def _same_lists(mode, ratio, soft="none", eps=0.0, G=1.3, ngroups=5, seed=0):
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(16, seed=8, rms=0.2)
    kw = dict(periodic=True, box=1.0, theta=0.4, order=4, soft=soft, eps=eps, G=G,
              merge_groups=mode, rcb_cell=512, ewald_ratio=ratio)
    print("periodic same lists mode=%d ratio=%g soft=%s" % (mode, ratio, soft), flush=True)
    plan = batch.Plan(pos, mass, **kw)
    out = {}
    for v in ("f64", "rel32c"):
        plan.evaluate(v)
        out[v] = plan.result(v)
    plan.close()
    H = batch.HostPlan(pos, mass, **kw)
    perm, spos = H.sorted()
    smass = mass[perm]
    for v in out:
        out[v] = out[v][perm]
    groups = np.random.default_rng(seed).choice(H.stats()["ngroups"], ngroups, replace=False)
    return H, spos, smass, out, [int(g) for g in groups]


@pytest.mark.parametrize("mode,ratio", [(1, 0.0), (5, 0.0), (1, 1.0), (5, 1.0)])
def test_periodic_same_lists(gpu_lib, mode, ratio):
    """ratio 0: every source's correction per pair; ratio 1: every source whose
    image no sink of the group sees differently through the group's expansion
    (the groups here are a fifth of the box: the expansion's own truncation is
    large, which the oracle's model shares).  f64 == oracle to 1e-9 of the rms
    force and potential; rel32c == f64 to 1e-5 of the rms force and 1e-6 of
    G sum(m) in the potential (FP32 sums of terms up to ~3 G sum(m) that cancel
    to ~1e-2 of it); the two ways differ (the comparison is not vacuous)."""
    G = 1.3
    H, spos, smass, out, groups = _same_lists(mode, ratio, G=G)
    rms_a = np.sqrt((out["f64"][:, :3] ** 2).sum(1).mean())
    rms_p = np.sqrt((out["f64"][:, 3] ** 2).mean())
    gap = 0.0
    for g in groups:
        _, ref = listoracle.group_reference_periodic(H, g, spos, smass, G=G, nodes=64,
                                                     ratio=ratio)
        slots = H.group_slots(g)
        got = out["f64"][slots]
        np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=0, atol=TOL * rms_a)
        np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=TOL * rms_p)
        r32 = out["rel32c"][slots]
        assert np.abs(r32[:, :3] - got[:, :3]).max() < 1e-5 * rms_a
        assert np.abs(r32[:, 3] - got[:, 3]).max() < 1e-6 * G * smass.sum()
        _, other = listoracle.group_reference_periodic(H, g, spos, smass, G=G, nodes=64,
                                                       ratio=1.0 - ratio)
        gap = max(gap, np.abs(other[:, :3] - ref[:, :3]).max())
    assert gap > 1e4 * TOL * rms_a, gap / rms_a


@pytest.mark.parametrize("ratio", [0.0, 1.0])
def test_periodic_same_lists_detects_planted_mistakes(gpu_lib, ratio):
    """The comparison above fails on a 1e-3 error in the correction or a flipped
    off-diagonal Jacobian sign (planted in the oracle, on the CPU): the f64
    result sits far outside the tolerance of every such reference."""
    G = 1.3
    H, spos, smass, out, groups = _same_lists(1, ratio, G=G, ngroups=2, seed=1)
    rms_a = np.sqrt((out["f64"][:, :3] ** 2).sum(1).mean())
    for g in groups:
        slots = H.group_slots(g)
        got = out["f64"][slots]
        for planted in ("scale_c", "flip_jac"):
            _, bad = listoracle.group_reference_periodic(H, g, spos, smass, G=G, nodes=64,
                                                         ratio=ratio, perturb=(planted,))
            gap = np.abs(got[:, :3] - bad[:, :3]).max()
            assert gap > 1e3 * TOL * rms_a, (planted, gap / rms_a)


def test_periodic_same_lists_plummer(gpu_lib):
    """Plummer-softened cells and bodies (the softened self pair keeps its
    -m/eps potential and the correction's limit psi(0)), both ways."""
    G = 1.0
    for ratio in (0.0, 1.0):
        H, spos, smass, out, groups = _same_lists(5, ratio, soft="plummer", eps=2e-3, G=G,
                                                  ngroups=3)
        rms_a = np.sqrt((out["f64"][:, :3] ** 2).sum(1).mean())
        rms_p = np.sqrt((out["f64"][:, 3] ** 2).mean())
        for g in groups:
            _, ref = listoracle.group_reference_periodic(H, g, spos, smass, soft="plummer",
                                                         eps=2e-3, G=G, nodes=64, ratio=ratio)
            got = out["f64"][H.group_slots(g)]
            np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=0, atol=TOL * rms_a)
            np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=TOL * rms_p)


def test_periodic_matches_ewald_direct_sum(gpu_lib):
    """Accelerations at the tree's own error against the exact lattice sum;
    potentials equal the reference's sum after its self constant is exchanged
    for the kernels' (m_i (2.8373 - pi/4) / L)."""
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(16, seed=8, rms=0.2)
    for mode in (1, 5):
        print("periodic vs Ewald direct sum mode=%d" % mode, flush=True)
        plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.4, order=4, soft="none",
                          eps=0.0, merge_groups=mode, rcb_cell=512)
        plan.evaluate("rel32c")
        acc = plan.result("rel32c")
        sample = np.random.default_rng(2).choice(pos.shape[0], 48, replace=False)
        direct = plan.direct(sample)  # FP64, nearest image + the FP64 table, self pair included
        plan.close()
        ref = oracle.direct_ewald(pos[sample], pos, mass, L=1.0)
        err = np.linalg.norm(acc[sample, :3] - ref[:, :3], axis=1)
        rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
        assert np.median(err) / rms < 1e-2, np.median(err) / rms
        assert err.max() / rms < 5e-2, err.max() / rms
        phi_ref = ref[:, 3] + mass[sample] * (XI - np.pi / 4)
        rms_phi = np.sqrt((phi_ref ** 2).mean())
        dphi = np.abs(acc[sample, 3] - phi_ref)
        # the correction's second order: its trace part is in the kernels, the
        # traceless part (~1e-2 of it) and the multipole truncation remain
        assert np.median(dphi) / rms_phi < 1e-2, np.median(dphi) / rms_phi
        # the library's FP64 direct sum: the same potentials to the table's accuracy
        np.testing.assert_allclose(direct[:, 3], phi_ref, rtol=0, atol=1e-4 * rms_phi)


def test_momentum_is_nearly_conserved(gpu_lib):
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(16, seed=9, rms=0.3)
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.5, order=4, soft="plummer",
                      eps=1e-3)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    net = np.abs((mass[:, None] * acc[:, :3]).sum(0)).max()
    typical = (mass[:, None] * np.abs(acc[:, :3])).sum()
    assert net < 1e-2 * typical
    plan.close()


def test_periodic_direct_sum_matches_exact_ewald(gpu_lib):
    """The GPU FP64 direct sum of a periodic plan (nearest image + the FP64
    Ewald table: nearest node plus its derivatives) against the oracle's exact
    lattice sum per pair: what remains is the table's interpolation error."""
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(12, seed=11, rms=0.3)
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.5, order=4, soft="none", eps=0.0)
    sample = np.random.default_rng(3).choice(pos.shape[0], 40, replace=False)
    got = plan.direct(sample)
    ref = oracle.direct_ewald(pos[sample], pos, mass, L=1.0)
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1)
    assert np.median(err) / rms < 1e-4, np.median(err) / rms
    assert err.max() / rms < 1e-3, err.max() / rms
    plan.close()


def test_update_keeps_a_crossing_body_beside_its_cell(gpu_lib):
    """A body that crosses the box boundary under a fixed tree (the caller hands it
    in wrapped into [0, L)) stays at the image next to its slot: forces match a
    plan built afresh from the new positions to the tree's accuracy, where a
    jump of a whole box inside its leaf would spoil every list holding that leaf."""
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(12, seed=12, rms=0.2)
    kw = dict(periodic=True, box=1.0, theta=0.5, order=4, soft="plummer", eps=1e-3)
    plan = batch.Plan(pos, mass, **kw)
    new = pos.copy()
    k = int(np.argmin(pos[:, 0]))
    new[k, 0] = (-0.5 * pos[k, 0]) % 1.0  # just across x = 0, wrapped to just below 1
    assert new[k, 0] > 0.9
    plan.update(new, mass)
    plan.evaluate("rel32c")
    got = plan.result("rel32c")
    plan.close()
    fresh = batch.Plan(new, mass, **kw)
    fresh.evaluate("rel32c")
    ref = fresh.result("rel32c")
    fresh.close()
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1) / rms
    assert err.max() < 5e-2, err.max()
    assert np.median(err) < 1e-2, np.median(err)
