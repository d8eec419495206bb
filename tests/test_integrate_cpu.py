"""The integrator's oracle (nln/integrate.c restated) on the CPU: known
answers, the cosmological step's reduction to the plain one, ConvertV's round
trip.  The GPU kernels are checked against this oracle in
test_integrate_gpu.py."""
import numpy as np

import oracle

BODY = np.dtype([("mass", "<f4"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("acc", "<f4", 3),
                 ("phi", "<f4"), ("nterms", "<f4"), ("ident", "<i8")])


def _table(n, seed):
    rng = np.random.default_rng(seed)
    t = np.zeros(n, BODY)
    t["mass"] = rng.random(n) + 0.1
    t["pos"] = rng.normal(size=(n, 3))
    t["vel"] = rng.normal(size=(n, 3))
    t["acc"] = rng.normal(size=(n, 3))
    t["phi"] = -rng.random(n)
    t["ident"] = np.arange(n)
    return t


def test_known_answer():
    """One body by hand: v' = v + dtv a, x' = x + dt v' (in FP64, rounded to
    float), ke at the synchronised velocity v + (tpos - tvel) a."""
    t = np.zeros(1, BODY)
    t["mass"] = 2.0
    t["pos"] = [1.0, 2.0, 3.0]
    t["vel"] = [0.5, -1.0, 0.25]
    t["acc"] = [1.0, 2.0, -4.0]
    t["phi"] = -3.0
    tp, tv, ke, pe = oracle.integrate(t, dt=0.5, dtv=0.25, tpos=1.0, tvel=0.75)
    assert (tp, tv) == (1.5, 1.0)
    vs = np.array([0.5, -1.0, 0.25]) + 0.25 * np.array([1.0, 2.0, -4.0])
    assert ke == 0.5 * 2.0 * (vs @ vs)
    assert pe == 0.5 * 2.0 * -3.0
    np.testing.assert_array_equal(t["vel"][0], np.float32([0.75, -0.5, -0.75]))
    np.testing.assert_array_equal(t["pos"][0], np.float32([1.375, 1.75, 2.625]))


def test_out_table_and_in_place_agree():
    a = _table(500, 1)
    b = a.copy()
    out = np.zeros(500, BODY)
    r1 = oracle.integrate(a, 0.01, 0.02, 3.0, 2.99, out=out)
    r2 = oracle.integrate(b, 0.01, 0.02, 3.0, 2.99)
    assert r1 == r2
    np.testing.assert_array_equal(out["pos"], b["pos"])
    np.testing.assert_array_equal(out["vel"], b["vel"])
    np.testing.assert_array_equal(a["pos"], _table(500, 1)["pos"])  # input untouched


def test_cosmo_step_with_unit_scale_factor_is_the_plain_step():
    """a0 = a1 = 1, kicks = dtv and the synchronising kick = tpos - tvel: the
    comoving conversions are exact and CosmoIntegrate reduces to Integrate."""
    a = _table(300, 2)
    b = a.copy()
    _, _, ke1, pe1 = oracle.integrate(a, 0.03, 0.05, 1.0, 0.98)
    _, _, ke2, pe2 = oracle.integrate(b, 0.03, 0.05, 1.0, 0.98,
                                      cosmo=(1.0, 1.0, 1.0 - 0.98, 0.05, 0.03))
    assert (ke1, pe1) == (ke2, pe2)
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["vel"], b["vel"])


def test_convert_v_round_trip():
    a = _table(200, 3)
    v0 = a["vel"].copy()
    oracle.convert_v(a, 4.0, 1)   # power-of-two scale: exact both ways
    np.testing.assert_array_equal(a["vel"], v0 / 4)
    oracle.convert_v(a, 4.0, 0)
    np.testing.assert_array_equal(a["vel"], v0)
