"""Periodic evaluation (nearest images + the 64^3 FP64 Ewald table) on the GPU.

Accelerations are compared with the oracle's FP64 Ewald direct sum
(oracle.direct_ewald, nln/ewald.c + nln/grav_n2.c pattern) on sampled bodies;
the error is measured against the rms acceleration because the forces in a
weakly perturbed lattice are small differences of large pair forces.  The
potentials are not compared: the reference's ewald() gives a self pair the
constant pi/(alpha^2 L^3) while a body's own image sum in the table path uses
the limit 2 x 1.4186487533 / L; conventions differ by a per-body constant.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("expand", [False, True])
def test_perfect_lattice_has_no_force(gpu_lib, expand):
    from pgrav import batch
    ng = 12
    g = (np.arange(ng) + 0.5) / ng
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([gx, gy, gz], -1).reshape(-1, 3)
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.5, order=4, soft="none", eps=0.0,
                      ewald_expand=expand)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    pair = mass[0] * ng ** 2  # one neighbour's pull at one lattice spacing
    # the exact answer is 0; what remains is the theta = 0.5 expansion error of the
    # cells (an FP64 CPU evaluation of the same lists with exact ewald() per pair
    # leaves 1.4% of a pair force)
    assert np.abs(acc[:, :3]).max() < 3e-2 * pair, np.abs(acc[:, :3]).max() / pair
    plan.close()


@pytest.mark.parametrize("mode,expand", [(1, False), (5, False), (1, True)])
def test_periodic_matches_ewald_direct_sum(gpu_lib, mode, expand):
    """mode 5: bisection groups (body sets) with the nearest image + Ewald table;
    expand: the cells' corrections expanded about the group centre, whose
    first-order error in the sink offset is large for groups this big relative to
    the box (radius ~0.1 L): about 2e-2 of the rms force here."""
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(16, seed=8, rms=0.2)
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.4, order=4, soft="none", eps=0.0,
                      merge_groups=mode, rcb_cell=512, ewald_expand=expand)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    sample = np.random.default_rng(2).choice(pos.shape[0], 48, replace=False)
    ref = oracle.direct_ewald(pos[sample], pos, mass, L=1.0)
    err = np.linalg.norm(acc[sample, :3] - ref[:, :3], axis=1)
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    tol = 4e-2 if expand else 1e-2
    assert np.median(err) / rms < tol, np.median(err) / rms
    assert err.max() / rms < 5 * tol, err.max() / rms
    # FP64 batched variant agrees with the FP32 one to FP32 precision
    plan.evaluate("f64")
    f64 = plan.result("f64")
    d = np.linalg.norm(f64[:, :3] - acc[:, :3], axis=1).max() / rms
    assert d < 1e-4, d
    plan.close()


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
    lattice sum per pair:
    what remains is the table's interpolation error."""
    from pgrav import batch, ics
    pos, mass = ics.lattice_grf(12, seed=11, rms=0.3)
    plan = batch.Plan(pos, mass, periodic=True, box=1.0, theta=0.5, order=4, soft="none", eps=0.0)
    sample = np.random.default_rng(3).choice(pos.shape[0], 40, replace=False)
    got = plan.direct(sample)
    ref = oracle.direct_ewald(pos[sample], pos, mass, L=1.0)
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1)
    assert np.median(err) / rms < 1e-3, np.median(err) / rms
    assert err.max() / rms < 1e-2, err.max() / rms
    plan.close()
