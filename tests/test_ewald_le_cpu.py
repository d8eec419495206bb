"""Explicit-image periodic scheme on the CPU: the far field's lattice
coefficients and polynomials against the reference's own constants
(nln/ewald_le2.c), the library's host expansion against the oracle's direct
pair sum, the whole scheme against the Ewald sum, and the image lists'
coverage (Nimage x N interactions per sink, nln/mac.c:734)."""
import itertools
import re

import numpy as np
import pytest

import listoracle
import oracle
from oracle import ewald_le
from pgrav import batch, ics
from pgrav._lib import PgravError

LE_XI = 2.837297479480620  # csrc/pgrav_le.h


@pytest.mark.parametrize("nimage", [1, 2])
def test_lattice_coefficients_match_reference_constants(nimage):
    """c_k = -t_k: the library's lattice sums reproduce the eight constants the
    reference hard-codes for each nimage (nln/ewald_le2.c:347-365) through its
    coefficient lines (:65, :111, :217)."""
    t = ewald_le.lattice_t(nimage)
    for k, tk in zip((4, 6, 8), t):
        np.testing.assert_allclose(batch.le_lattice(k, nimage), -tk, rtol=1e-10, atol=1e-16)


def test_cubic_harmonics():
    """K_k: harmonic (the Laplacian of the monomial sum vanishes term by term),
    invariant under axis permutations and reflections, with the coefficients of
    the reference's Q[1], Q[14], Q[31] (fixture)."""
    gold = ewald_le.constants()["harmonic_coefficients"]
    rng = np.random.default_rng(0)
    r = rng.normal(size=(50, 3))
    for k, terms in ewald_le.K.items():
        lap = {}
        for (a, b, c), v in terms.items():
            for ax, e in enumerate((a, b, c)):
                if e >= 2:
                    m = [a, b, c]
                    m[ax] -= 2
                    lap[tuple(m)] = lap.get(tuple(m), 0.0) + v * e * (e - 1)
        assert all(abs(v) < 1e-12 for v in lap.values()), k
        v, _ = ewald_le.poly(terms, r)
        for p in itertools.permutations(range(3)):
            np.testing.assert_allclose(ewald_le.poly(terms, r[:, p])[0], v, rtol=1e-12)
        np.testing.assert_allclose(ewald_le.poly(terms, r * [-1, 1, -1])[0], v, rtol=1e-12)
        for name, c in gold[str(k)].items():
            e = dict.fromkeys("xyz", 0)
            e.update({ax: int(d) for ax, d in re.findall(r"([xyz])(\d+)", name)})
            assert terms.get((e["x"], e["y"], e["z"]), 0.0) == c, (k, name)


@pytest.mark.parametrize("nimage", [1, 2])
def test_host_far_field_matches_oracle(nimage):
    """The library's moment expansion (host/ewald_le.cpp: global moments, W)
    against the oracle's direct pair sum of the reference's polynomials."""
    rng = np.random.default_rng(nimage)
    pos = rng.random((60, 3)) * 2.0
    mass = rng.random(60) + 0.1
    ref = ewald_le.far_field(pos, mass, box=2.0, nimage=nimage, G=0.7)
    got = batch.ewald_le_host(pos, mass, box=2.0, nimage=nimage, G=0.7)
    scale = np.abs(ref[:, :3]).max()
    np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=0, atol=1e-10 * scale)
    # the potentials differ only by the self constant's digits: 2 x 1.4186487533
    # in the reference, 2.8372974794806 here
    dxi = (ewald_le.XI_REF - LE_XI) * 0.7 / 2.0 * (mass.sum() - mass)
    np.testing.assert_allclose(got[:, 3] + dxi, ref[:, 3], rtol=1e-11)


def test_self_constant():
    """xi: the Ewald potential of nln/ewald.c (oracle) near the origin is
    -1/r + xi - (2 pi / 3) r^2 + O(r^4)."""
    x = np.array([2e-3, -1e-3, 1.5e-3])
    r = np.linalg.norm(x)
    _, phi = oracle.ewald(x)
    assert abs(phi + 1 / r + 2 * np.pi / 3 * r * r - LE_XI) < 1e-9
    assert abs(LE_XI - ewald_le.XI_REF) < 3e-8


@pytest.mark.parametrize("nimage,tol", [(1, 1e-3), (2, 2e-6)])
def test_images_plus_far_field_is_the_ewald_sum(nimage, tol):
    """Near images summed directly + the far field == the Ewald sum of
    nln/ewald.c (oracle) up to the expansion's truncation at degree 8; the
    potential convention leaves out each body's own images (the m_i term of
    nln/ewald_le2.c:44), i.e. the Ewald pair sum minus its self pair."""
    rng = np.random.default_rng(10 + nimage)
    n = 40
    pos = rng.random((n, 3))
    mass = rng.random(n) + 0.5
    tot = ewald_le.near_images(pos, mass, 1.0, nimage) + batch.ewald_le_host(pos, mass, 1.0, nimage)
    ex = oracle.direct_ewald(pos, pos, mass, 1.0)
    ex[:, 3] -= mass * np.pi / 4.0  # the self pair of nln/ewald.c: pi / (alpha^2 L^3), alpha = 2
    rms = np.sqrt((ex[:, :3] ** 2).sum(1).mean())
    assert np.linalg.norm(tot[:, :3] - ex[:, :3], axis=1).max() < tol * rms
    assert np.abs(tot[:, 3] - ex[:, 3]).max() < 10 * tol * np.abs(ex[:, 3]).mean()


@pytest.mark.parametrize("nimage", [1, 2])
def test_host_far_field_sb_matches_oracle(nimage):
    """The subtract_background form: the library's moment expansion with the cube
    folded into W against the oracle's pair sum less the cube's integral."""
    rng = np.random.default_rng(20 + nimage)
    pos = rng.random((50, 3)) * 2.0
    mass = rng.random(50) + 0.1
    ref = ewald_le.far_field_sb(pos, mass, box=2.0, nimage=nimage, G=0.7)
    got = batch.ewald_le_host(pos, mass, box=2.0, nimage=nimage, G=0.7, subtract_background=True)
    scale = np.abs(ref[:, :3]).max()
    np.testing.assert_allclose(got[:, :3], ref[:, :3], rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(got[:, 3], ref[:, 3], rtol=0, atol=1e-10 * np.abs(ref[:, 3]).max())


@pytest.mark.parametrize("nimage,tol", [(1, 1e-3), (2, 2e-6)])
def test_images_less_background_plus_sb_far_field_is_the_ewald_sum(nimage, tol):
    """The subtract_background scheme's accounting: the explicit images of the
    bodies, less the uniform background over the same (2 nimage + 1)^3 boxes (a
    cube of density -M / L^3, its field in closed form), plus the sb far field, is
    the Ewald sum (nln/ewald.c, oracle) -- the dipole term turns the cubic
    summation into the Ewald (tin-foil) convention."""
    from oracle import sbmac
    rng = np.random.default_rng(30 + nimage)
    n = 40
    pos = rng.random((n, 3))
    pos[:8] *= 0.3  # a net dipole about the box centre
    mass = rng.random(n) + 0.5
    M = mass.sum()
    half = (2 * nimage + 1) * 0.5
    bg = np.array([-M * sbmac.cube_field(p - 0.5, half) for p in pos])
    far = batch.ewald_le_host(pos, mass, 1.0, nimage, subtract_background=True)
    tot = ewald_le.near_images(pos, mass, 1.0, nimage)[:, :3] + bg + far[:, :3]
    ex = oracle.direct_ewald(pos, pos, mass, 1.0)
    rms = np.sqrt((ex[:, :3] ** 2).sum(1).mean())
    assert np.linalg.norm(tot - ex[:, :3], axis=1).max() < tol * rms
    # without the dipole term the sum misses (4 pi / 3) G P / L^3
    dip = 4 * np.pi / 3 * (mass[:, None] * (pos - 0.5)).sum(0)
    assert np.linalg.norm(tot + dip - ex[:, :3], axis=1).max() > 10 * tol * rms


@pytest.mark.parametrize("nimage", [0, 1])
def test_image_lists_cover_n_per_image(nimage):
    """Each sink meets every body once per explicit image (nln/mac.c:734:
    interactions == Nimage * N)."""
    pos, mass = ics.lattice_grf(12, seed=3)
    n = pos.shape[0]
    H = batch.HostPlan(pos, mass, periodic="images", nimage=nimage, box=1.0, theta=0.6, order=4,
                       eps=1e-3)
    f = 2 * nimage + 1
    st = H.stats()
    for g in range(0, st["ngroups"], 7):
        per = np.zeros(f ** 3, np.int64)
        for kind in listoracle.ORDER:
            for e, im in zip(H.list(g, kind), H.list_images(g, kind)):
                per[im] += H.cell(int(e))["count"]
        np.add.at(per, H.list_images(g, "body").astype(np.int64), 1)
        assert (per == n).all()


def test_image_lists_and_far_field_reproduce_the_ewald_sum():
    """The host's image lists, evaluated by the oracle (cells rebuilt from their
    bodies at the image position), plus the far field == the Ewald sum of
    nln/ewald.c for the sinks of sampled groups: what remains is the order-4
    expansion error at theta = 0.4."""
    pos, mass = ics.lattice_grf(10, seed=5, rms=0.3)
    H = batch.HostPlan(pos, mass, periodic="images", nimage=1, box=1.0, theta=0.4, order=4,
                       soft="none", eps=0.0)
    perm, spos = H.sorted()
    smass = mass[perm]
    far = batch.ewald_le_host(pos, mass, box=1.0, nimage=1)
    st = H.stats()
    errs, rmss = [], []
    for g in range(0, st["ngroups"], max(1, st["ngroups"] // 6)):
        first, acc = listoracle.group_reference(H, g, spos, smass, soft="none", eps=0.0,
                                                images=(1.0, 1))
        idx = perm[first:first + acc.shape[0]]
        acc = acc + far[idx]
        ex = oracle.direct_ewald(pos[idx], pos, mass, 1.0)
        errs.append(np.linalg.norm(acc[:, :3] - ex[:, :3], axis=1))
        rmss.append((ex[:, :3] ** 2).sum(1))
        # potentials: the lattice's is a small difference of O(G M / L) terms
        dphi = acc[:, 3] - (ex[:, 3] - mass[idx] * np.pi / 4.0)
        assert np.abs(dphi).max() < 1e-3 * mass.sum()
    err = np.concatenate(errs)
    rms = np.sqrt(np.concatenate(rmss).mean())
    assert np.median(err) < 5e-3 * rms and err.max() < 3e-2 * rms, (np.median(err) / rms,
                                                                      err.max() / rms)


def test_image_parameters_are_checked():
    pos, mass = ics.lattice_grf(6, seed=1)
    with pytest.raises(PgravError):
        batch.HostPlan(pos, mass, periodic="images", nimage=3, box=1.0)
    with pytest.raises(ValueError):
        batch.make_params(periodic="mirror")
