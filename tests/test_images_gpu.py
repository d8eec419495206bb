"""Explicit-image periodic evaluation on the GPU (periodic = 2): the walk's
(2 nimage + 1)^3 images in the lists (nln/mac.c:91-111) plus the far images
and background by the local expansion (nln/main.c:1811-1862,
nln/ewald_le2.c), the reference's production periodic scheme.

* FP64 variant == the oracle's FP64 evaluation of the same lists (cells
  rebuilt from their bodies at the image position) + the host far field, to
  round-off: pins the image decoding in the kernels and the far-field kernels.
* FP32 variant vs the exact Ewald sum (the plan's FP64 direct sum with the
  Ewald table): the order-4 expansion error, like the nearest-image mode.
* A captured graph replays the whole evaluation (tree kernel + the three
  far-field kernels) with identical results.
"""
import numpy as np
import pytest

import listoracle

pytestmark = pytest.mark.gpu


def _lattice(ng, seed, rms=0.3):
    from pgrav import ics
    return ics.lattice_grf(ng, seed=seed, rms=rms)


@pytest.mark.parametrize("nimage,mode", [(1, 1), (2, 1), (1, 5)])
def test_fp64_images_match_oracle_lists(gpu_lib, nimage, mode):
    """mode 5: bisection groups (body sets) under the explicit images."""
    from pgrav import batch
    pos, mass = _lattice(10, seed=21)
    kw = dict(periodic="images", nimage=nimage, box=1.0, theta=0.5, order=4, soft="none",
              eps=0.0, G=1.3, merge_groups=mode, rcb_cell=256)
    print("images nimage=%d mode=%d n=%d" % (nimage, mode, pos.shape[0]), flush=True)
    plan = batch.Plan(pos, mass, **kw)
    plan.evaluate("f64")
    got = plan.result("f64")
    plan.close()
    H = batch.HostPlan(pos, mass, **kw)
    perm, spos = H.sorted()
    far = batch.ewald_le_host(pos, mass, box=1.0, nimage=nimage, G=1.3)
    st = H.stats()
    scale = np.abs(far[:, :3]).max() + np.abs(got[:, :3]).max()
    for g in range(0, st["ngroups"], max(1, st["ngroups"] // 8)):
        _, ref = listoracle.group_reference(H, g, spos, mass[perm], soft="none", eps=0.0,
                                            G=1.3, images=(1.0, nimage))
        idx = perm[H.group_slots(g)]
        ref = ref + far[idx]
        np.testing.assert_allclose(got[idx, :3], ref[:, :3], rtol=0, atol=1e-10 * scale)
        np.testing.assert_allclose(got[idx, 3], ref[:, 3], rtol=0, atol=1e-10 * mass.sum() * 1.3)


def test_fp32_images_match_exact_ewald(gpu_lib):
    from pgrav import batch
    pos, mass = _lattice(16, seed=8, rms=0.2)
    print("images rel32c vs Ewald n=%d" % pos.shape[0], flush=True)
    plan = batch.Plan(pos, mass, periodic="images", nimage=2, box=1.0, theta=0.4, order=4,
                      soft="none", eps=0.0)
    plan.evaluate("rel32c")
    acc = plan.result("rel32c")
    sample = np.random.default_rng(2).choice(pos.shape[0], 64, replace=False)
    ref = plan.direct(sample)  # FP64: nearest image + Ewald table, self pair left out
    rms = np.sqrt((ref[:, :3] ** 2).sum(1).mean())
    err = np.linalg.norm(acc[sample, :3] - ref[:, :3], axis=1)
    assert np.median(err) / rms < 1e-2, np.median(err) / rms
    assert err.max() / rms < 5e-2, err.max() / rms
    dphi = np.abs(acc[sample, 3] - ref[:, 3])
    assert dphi.max() < 1e-3 * mass.sum(), dphi.max()
    plan.close()


def test_images_graph_replay(gpu_lib):
    from pgrav import batch
    pos, mass = _lattice(12, seed=4)
    print("images graph replay", flush=True)
    plan = batch.Plan(pos, mass, periodic="images", nimage=1, box=1.0, theta=0.6, order=4,
                      soft="plummer", eps=1e-3)
    plan.evaluate("rel32c")
    eager = plan.result("rel32c")
    plan.graph_capture("rel32c")
    for _ in range(3):
        plan.graph_launch()
        plan.synchronize()
        np.testing.assert_array_equal(plan.result("rel32c"), eager)
        plan.evaluate("rel32c")  # eager work between the replays
        plan.synchronize()
    plan.close()


@pytest.mark.parametrize("sb", [0, 1])
@pytest.mark.parametrize("nimage", [1, 2])
def test_far_field_device_matches_host(gpu_lib, nimage, sb):
    """pgrav_ewald_le_device on caller device arrays (both forms: FixCubeEwaldLE with
    and without subtract_background) equals the host FP64 expansion, which the CPU
    tests pin to the oracle's pair sum and to the Ewald sum."""
    import torch
    from pgrav import batch
    rng = np.random.default_rng(40 + nimage + 2 * sb)
    n = 5000
    pos = rng.random((n, 3)) * 2.0
    pos[:500] *= 0.2  # a dipole about the box centre
    mass = rng.random(n) + 0.5
    want = batch.ewald_le_host(pos, mass, box=2.0, nimage=nimage, G=0.7,
                               subtract_background=bool(sb))
    dp = torch.from_numpy(pos).cuda()
    dm = torch.from_numpy(mass).cuda()
    base = torch.full((n, 4), 0.25, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    batch.ewald_le_device(dp.data_ptr(), dm.data_ptr(), n, base.data_ptr(), box=2.0, nimage=nimage,
                          G=0.7, subtract_background=bool(sb))
    torch.cuda.synchronize()
    got = base.cpu().numpy() - 0.25  # it adds into acc4
    scale = np.abs(want).max(axis=0)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-11 * scale.max())
