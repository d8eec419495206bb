"""Seeded synthetic particle sets for the measured workloads.

All sets have equal-mass bodies of total mass 1 (G = 1 units).  The
reference's own generator (nln/main.c:444-495, do_test) only makes uniform or
centrally condensed spheres/cubes, so these are written from the workload
definitions:

plummer(n)        Plummer sphere, scale radius a, truncated at rmax = 10 a.
clustered(n)      nhalo Plummer halos with halo masses dN/dM ~ M^-alpha over
                  [M0, mrange M0], centres uniform in the unit cube, scale radius
                  a_h = a_max (M_h / M_max)^(1/3) with a_max = 0.02, each truncated
                  at 10 a_h; halo bodies are shared in proportion to halo mass;
                  plus a fraction `background` uniform in the unit cube.
lattice_grf(ng)   ng^3 unit-box lattice displaced by a Gaussian random field whose
                  three components each have P(k) ~ k^slope, scaled to an rms
                  displacement of `rms` grid spacings; periodic, wrapped to [0, 1).
"""
import numpy as np


def _plummer_radii(rng, n, a, rmax_over_a=10.0):
    # r = a (X^-2/3 - 1)^-1/2; r <= rmax  <=>  X <= (1 + (a/rmax)^2)^-3/2
    xmax = (1.0 + 1.0 / rmax_over_a ** 2) ** -1.5
    x = rng.uniform(0.0, xmax, n)
    x = np.maximum(x, 1e-300)
    return a / np.sqrt(x ** (-2.0 / 3.0) - 1.0)


def _iso(rng, n):
    u = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * np.pi, n)
    s = np.sqrt(1.0 - u * u)
    return np.stack([s * np.cos(phi), s * np.sin(phi), u], axis=1)


def plummer(n, seed=1, a=1.0, rmax_over_a=10.0):
    rng = np.random.default_rng(seed)
    r = _plummer_radii(rng, n, a, rmax_over_a)
    pos = _iso(rng, n) * r[:, None]
    return pos, np.full(n, 1.0 / n)


def clustered(n, seed=2, nhalo=512, alpha=1.9, mrange=1000.0, background=0.2, a_max=0.02):
    rng = np.random.default_rng(seed)
    # halo masses by inverse CDF of dN/dM ~ M^-alpha on [1, mrange]
    e = 1.0 - alpha
    u = rng.uniform(0.0, 1.0, nhalo)
    mh = (1.0 + u * (mrange ** e - 1.0)) ** (1.0 / e)
    n_bg = int(round(background * n))
    n_h = n - n_bg
    share = mh / mh.sum() * n_h
    counts = np.floor(share).astype(np.int64)
    rest = n_h - counts.sum()
    counts[np.argsort(share - counts)[::-1][:rest]] += 1
    centres = rng.uniform(0.0, 1.0, (nhalo, 3))
    ah = a_max * (mh / mh.max()) ** (1.0 / 3.0)
    pos = np.empty((n, 3), np.float64)
    o = 0
    for h in range(nhalo):
        c = counts[h]
        if c == 0:
            continue
        r = _plummer_radii(rng, c, ah[h])
        pos[o:o + c] = centres[h] + _iso(rng, c) * r[:, None]
        o += c
    pos[o:] = rng.uniform(0.0, 1.0, (n - o, 3))
    perm = rng.permutation(n)  # input order carries no structure
    return pos[perm], np.full(n, 1.0 / n)


def lattice_grf(ng, seed=4, rms=0.2, slope=-2.0):
    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(ng) * ng
    kz = np.fft.rfftfreq(ng) * ng
    k2 = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = (k2 ** (slope / 4.0)).astype(np.float32)  # sqrt(P(k)), P ~ k^slope
    amp[0, 0, 0] = 0.0
    disp = np.empty((3, ng, ng, ng), np.float32)
    for c in range(3):
        w = rng.standard_normal((ng, ng, ng)).astype(np.float32)
        f = np.fft.rfftn(w)
        f *= amp
        disp[c] = np.fft.irfftn(f, s=(ng, ng, ng), axes=(0, 1, 2)).astype(np.float32)
    scale = rms / np.sqrt((disp.astype(np.float64) ** 2).sum(0).mean())
    g = (np.arange(ng) + 0.5) / ng
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([gx, gy, gz], axis=-1).reshape(-1, 3)
    pos += (disp.reshape(3, -1).T.astype(np.float64) * scale) / ng
    pos -= np.floor(pos)
    return pos, np.full(pos.shape[0], 1.0 / pos.shape[0])
