"""The integrator kernels (nln/integrate.c on the GPU) against the oracle:
positions and velocities bit for bit (the same FP64 operations, rounded as
the C source rounds them), energies to the summation order; the host-pointer
drop-in and the device entry points; a separate output table and in place;
and a few leapfrog steps of a Plummer sphere through the batched force
evaluation conserving energy."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

OUTBODY = np.dtype([("mass", "<f4"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("acc", "<f4", 3),
                    ("phi", "<f4"), ("ident", "<i8")])  # nln/physics.h:49-58, SAVE_ACC


def _table(n, seed):
    from pgrav.integrate import BODY
    rng = np.random.default_rng(seed)
    t = np.zeros(n, BODY)
    t["mass"] = rng.random(n) + 0.1
    t["pos"] = rng.normal(size=(n, 3))
    t["vel"] = rng.normal(size=(n, 3))
    t["acc"] = rng.normal(size=(n, 3)) * 10
    t["phi"] = -rng.random(n)
    t["ident"] = np.arange(n)
    return t


def _same(a, b, fields=("pos", "vel")):
    for f in fields:
        np.testing.assert_array_equal(a[f], b[f])


def test_host_dropin_matches_oracle(gpu_lib):
    from pgrav import integrate
    print("integrate host drop-in", flush=True)
    for n in (1, 255, 256, 257, 50000):
        a = _table(n, n)
        b = a.copy()
        r_gpu = integrate.integrate(a, 0.013, 0.027, 5.0, 4.987)
        r_ref = oracle.integrate(b, 0.013, 0.027, 5.0, 4.987)
        _same(a, b)
        assert r_gpu[:2] == r_ref[:2]
        np.testing.assert_allclose(r_gpu[2:], r_ref[2:], rtol=1e-13)
        np.testing.assert_array_equal(a["ident"], np.arange(n))  # other fields kept


def test_host_dropin_separate_output_table(gpu_lib):
    from pgrav import integrate
    a = _table(3000, 7)
    out = np.zeros(3000, OUTBODY)
    out["ident"] = 99
    ref_out = out.copy()
    integrate.integrate(a, 0.1, 0.2, 0.0, 0.0, out=out)
    oracle.integrate(a.copy(), 0.1, 0.2, 0.0, 0.0, out=ref_out)
    _same(out, ref_out)
    assert (out["ident"] == 99).all() and (out["mass"] == 0).all()


def test_device_entry_points_match_oracle(gpu_lib):
    import torch
    from pgrav import integrate
    from pgrav.integrate import BODY
    print("integrate device", flush=True)
    a = _table(10000, 3)
    dev = torch.from_numpy(a.view(np.uint8).copy()).cuda()
    tp, tv, ke, pe = integrate.integrate_device(dev.data_ptr(), a.shape[0], BODY, 0.01, 0.02, 2.0,
                                                1.99)
    got = dev.cpu().numpy().view(BODY)
    ref = a.copy()
    r = oracle.integrate(ref, 0.01, 0.02, 2.0, 1.99)
    _same(got, ref)
    assert (tp, tv) == r[:2]
    np.testing.assert_allclose((ke, pe), r[2:], rtol=1e-13)
    # CosmoIntegrate with its factors, then ConvertV both ways
    fac = (0.5, 0.52, 0.011, 0.021, 0.033)
    ke, pe = integrate.cosmo_integrate_device(dev.data_ptr(), a.shape[0], BODY, *fac)
    r = oracle.integrate(ref, 0, 0, 0, 0, cosmo=fac)
    _same(dev.cpu().numpy().view(BODY), ref)
    np.testing.assert_allclose((ke, pe), r[2:], rtol=1e-13)
    for scale, phys in ((0.37, 1), (1.9, 0)):
        integrate.convert_v_device(dev.data_ptr(), a.shape[0], BODY, scale, phys)
        oracle.convert_v(ref, scale, phys)
        _same(dev.cpu().numpy().view(BODY), ref, ("vel",))
    torch.cuda.synchronize()


def test_bad_arguments_raise(gpu_lib):
    from pgrav import integrate
    from pgrav._lib import PgravError
    a = _table(10, 1)
    with pytest.raises(PgravError):
        integrate.integrate_device(0, 10, a.dtype, 0.1, 0.1, 0.0, 0.0)


def test_leapfrog_conserves_energy(gpu_lib):
    """Kick-drift-kick steps of a Plummer sphere: the batched force evaluation
    (FP64 variant, theta = 0.3) and the integrator; total energy drifts by far
    less than a per cent over 20 steps of 1/200 of a crossing time."""
    from pgrav import batch, ics, integrate
    print("leapfrog Plummer", flush=True)
    pos, mass = ics.plummer(4096, seed=5)
    n = pos.shape[0]
    t = np.zeros(n, integrate.BODY)
    t["mass"] = mass
    t["pos"] = pos
    t["ident"] = np.arange(n)
    eps = 0.01
    plan = batch.Plan(t["pos"].astype(np.float64), mass, theta=0.3, order=4, soft="plummer",
                      eps=eps)

    def forces():
        plan.update(t["pos"].astype(np.float64), mass)
        plan.evaluate("f64")
        r = plan.result("f64")
        t["acc"] = r[:, :3]
        t["phi"] = r[:, 3] + mass / eps  # the softened self term
    forces()
    dt = 0.005
    energies = []
    tpos = tvel = 0.0
    # start the leapfrog: velocities half a step behind (tvel = tpos - dt/2)
    tpos, tvel, _, _ = integrate.integrate(t, 0.0, -0.5 * dt, tpos, tvel)
    for step in range(20):
        tpos, tvel, ke, pe = integrate.integrate(t, dt, dt, tpos, tvel)
        energies.append(ke + pe)
        forces()
    plan.close()
    e = np.array(energies)
    assert np.abs(e - e[0]).max() < 2e-3 * abs(e[0]), (e[0], np.abs(e - e[0]).max())


def test_run_on_an_sdf_snapshot(gpu_lib, tmp_path):
    """pgrav.run: SDF in, steps of the reference's main loop (forces at tpos,
    Integrate with dtv = tpos + dt/2 - tvel), SDF out with accelerations; the
    same steps through the API give the same table."""
    from pgrav import ics, run, sdf
    print("run on SDF", flush=True)
    pos, mass = ics.plummer(2048, seed=9)
    recs = np.zeros(pos.shape[0], sdf.OUTBODY)
    recs["mass"] = mass
    recs["x"], recs["y"], recs["z"] = pos.T
    recs["ident"] = np.arange(pos.shape[0]) + 100
    src = tmp_path / "ic.sdf"
    sdf.write(src, recs, sdf.OUTBODY_DESC, sdf.output_header(pos.shape[0]))
    dst = tmp_path / "out.sdf"
    run.main([str(src), str(dst), "--steps", "3", "--dt", "0.01", "--eps", "0.01",
              "--theta", "0.4", "--rebuild", "2"])
    h, out = sdf.read(dst)
    assert h["iter"] == 3 and abs(h["tpos"] - 0.03) < 1e-15 and abs(h["tvel"] - 0.025) < 1e-15
    np.testing.assert_array_equal(out["ident"], recs["ident"])
    st = run.Stepper(recs, theta=0.4, eps=0.01, rebuild=2)
    for _ in range(3):
        st.step(0.01)
    st.forces()
    ref = st.records()
    st.close()
    for f in ("x", "y", "z", "vx", "vy", "vz", "ax", "phi"):
        np.testing.assert_array_equal(out[f], ref[f])
    assert h["ke"] > 0 and h["pe"] < 0
