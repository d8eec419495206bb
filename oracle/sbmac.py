"""TEST INFRASTRUCTURE ONLY -- restatement of the reference's subtract_background scheme.

Pieces, each with the reference lines it follows:

  mpole_add_cube(a, rho, raw)   nln/cofm.c:626-648: a cube of edge a and density rho
                                added to raw moments about its centre: m += a^3 rho,
                                x2 += a^5 rho / 12, x4 += a^7 rho / 80, x2y2 += a^7 rho / 144
                                (each axis / pair alike).
  cube_field(r, a)              the attraction at r of the cube [-a, a]^3 of unit density,
                                unit G: what nln/mac.c:688-694 subtracts at the leaf times
                                rho0 (cubic_acc, nln/ewald_le2.c:773, evaluates the same
                                field in FP32).  Closed form (the prism formula, Nagy 1966).
  cube_field_quad(r, a)         the same field by one analytic integration (asinh) and
                                adaptive quadrature over the other face coordinate: an
                                independent route used to pin cube_field.
  ucell(r0, rho0)               nln/mac.c:237-250: half-size and mass of a level's cell.
  dl_mac_sb(...)                one DLRcritMACsb call, nln/mac.c:1565-1771 (results,
                                background flags out, appends, near count), FP32 as the
                                reference's locals.

Parity: the arithmetic that decides (distances, radii, cube centres) is FP32 in the
order the reference writes it; the closed forms are FP64 (the reference's cubic_acc
is FP32, so its results differ from these by FP32 rounding).
"""
import math

import numpy as np
from scipy import integrate

ACCEPT, SPLIT_SINK, SPLIT_SRC = 0, 1, 2
CHUBITS = 42
# append codes (the library's call log)
M, Q, H, QB, HB, ANTI_BODY, ANTI_CELL, ANTI_EMPTY, MM = 0, 1, 2, 3, 4, 5, 6, 8, 20

f32 = np.float32


def mpole_add_cube(a, rho, raw):
    """raw: dict {(i, j, k): moment}; returns a new dict with the cube added."""
    out = dict(raw)
    a3 = a * a * a * rho
    a5 = a * a * a3
    a7 = a * a * a5

    def add(e, v):
        out[e] = out.get(e, 0.0) + v
    add((0, 0, 0), a3)
    for e2, e4 in (((2, 0, 0), (4, 0, 0)), ((0, 2, 0), (0, 4, 0)), ((0, 0, 2), (0, 0, 4))):
        add(e2, a5 / 12.0)
        add(e4, a7 / 80.0)
    for e in ((2, 2, 0), (2, 0, 2), (0, 2, 2)):
        add(e, a7 / 144.0)
    return out


def _face(c, y, z):
    """y ln(z + R) + z ln(y + R) - c atan(y z / (c R)): the antiderivative in y and z
    of 1 / sqrt(c^2 + y^2 + z^2)."""
    R = math.sqrt(c * c + y * y + z * z)

    def ylog(p, q):
        if p == 0.0:
            return 0.0
        t = q + R if q >= 0 else (c * c + p * p) / (R - q)
        return p * math.log(t)
    f = ylog(y, z) + ylog(z, y)
    if c != 0.0:
        f -= c * math.atan(y * z / (c * R))
    return f


def cube_field(r, a):
    """Attraction (unit G) at r of the uniform unit-density cube [-a, a]^3."""
    r = [float(v) for v in r]
    lo = [-a - v for v in r]
    hi = [a - v for v in r]
    g = np.zeros(3)
    for ax in range(3):
        b, c = (ax + 1) % 3, (ax + 2) % 3
        s = 0.0
        # int u / |u|^3 du over [lo, hi] = 1 / |u(lo)| - 1 / |u(hi)|
        for u, sign in ((lo[ax], 1.0), (hi[ax], -1.0)):
            s += sign * (_face(u, hi[b], hi[c]) - _face(u, hi[b], lo[c])
                         - _face(u, lo[b], hi[c]) + _face(u, lo[b], lo[c]))
        g[ax] = s
    return g


def cube_field_quad(r, a):
    """The same field: for each face u_ax = const the integral of 1/|u| over the face
    is done in z analytically (asinh) and in y by adaptive quadrature."""
    r = [float(v) for v in r]
    g = np.zeros(3)
    for ax in range(3):
        b, c = (ax + 1) % 3, (ax + 2) % 3
        ylo, yhi = -a - r[b], a - r[b]
        zlo, zhi = -a - r[c], a - r[c]
        s = 0.0
        for u, sign in ((-a - r[ax], 1.0), (a - r[ax], -1.0)):
            def fy(y, u=u):
                h = math.hypot(u, y)
                if h == 0.0:
                    return 0.0  # a set of measure zero
                return math.asinh(zhi / h) - math.asinh(zlo / h)
            pts = [0.0] if ylo < 0.0 < yhi else None
            v, _ = integrate.quad(fy, ylo, yhi, points=pts, epsabs=1e-13, epsrel=1e-12,
                                  limit=200)
            s += sign * v
        g[ax] = s
    return g


def ucell(r0, rho0, levels=CHUBITS):
    """[(halfsz, mass)] per level in FP32 (nln/mac.c:237-250, expand_root 0)."""
    a1 = f32(r0)
    e2 = f32(f32(2.0) * a1)
    a3 = f32(float(rho0) * float(f32(f32(e2 * e2) * e2)))
    out = []
    for _ in range(levels):
        out.append((a1, a3))
        a1 = f32(a1 * f32(0.5))
        a3 = f32(a3 * f32(0.125))
    return out


def _d2(a, b):
    d = [f32(f32(a[k]) - f32(b[k])) for k in range(3)]
    return f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))


def _outside(m, cen, cr):
    return any(abs(f32(f32(m[k]) - f32(cen[k]))) > f32(cr) for k in range(3))


def _cube_centre(corner, edge, off):
    return [f32(f32(corner[k]) + f32(f32(f32(0.5) * f32(edge)) + f32(off[k]))) for k in range(3)]


def dl_mac_sb(sink, sources, cells, node_info, node_geom, body_leaf, bodies, offsets, uc,
              rho0, box, dlfac, p2cut, p4cut, leaf_max_n, do_pM):
    """One DLRcritMACsb call.

    sink: dict pos, bmax, daughters, cen, cr, cr2, clevel, done;
    sources: [(entity, image, background flag)];
    cells (nnodes, 8) float32 mass, x, y, z, rcrit, bmax, rcrit_q, rcrit_h;
    node_info (nnodes, 4) first body, bodies, has children, level;
    node_geom (nnodes, 5) key-cube corner, edge, octant mask; body_leaf: leaf node per body;
    uc: ucell(r0, rho0).  Returns (results, flags out, appends [(code, entity, image)],
    near increments, interactions)."""
    results, flags, appends = [], [], []
    near, inter = 0, 0
    if sink["done"]:
        return ([SPLIT_SINK] * len(sources), [int(s[2]) for s in sources], [], 0, 0)
    sp = sink["pos"]
    cen = sink["cen"]
    cr, cr2 = f32(sink["cr"]), f32(sink["cr2"])
    clevel = int(sink["clevel"])
    central = offsets.shape[0] >> 1
    for ent, img, bs in sources:
        ent, img, bs = int(ent), int(img), bool(bs)
        off = offsets[img]
        if ent >= 0:
            c = cells[ent]
            r = [f32(f32(c[1 + k]) + f32(off[k])) for k in range(3)]
            nb = int(node_info[ent, 1])
            lev = int(node_info[ent, 3])
            rc, rq, rh = f32(c[4]), f32(c[6]), f32(c[7])
            isquad = p2cut != 0 and nb >= p2cut
            ishexa = p4cut != 0 and nb >= p4cut
            smallest = rh if ishexa else (rq if isquad else rc)
            if clevel != CHUBITS and lev < clevel:
                dr2 = _d2(r, cen)
                bmax = f32(f32(1.2) * cr)
            else:
                dr2 = _d2(r, sp)
                bmax = f32(sink["bmax"])
            if not bs and lev == clevel and _d2(r, cen) < cr2:
                bs = True
                near += 1
            if isquad and dr2 > f32(f32(rq + bmax) * f32(rq + bmax)):
                appends.append((QB if bs else Q, ent, img))
                inter += nb
                res = ACCEPT
            elif ishexa and dr2 > f32(f32(rh + bmax) * f32(rh + bmax)):
                appends.append((HB if bs else H, ent, img))
                inter += nb
                res = ACCEPT
            else:
                res = (SPLIT_SINK if f32(f32(dlfac) * f32(sink["bmax"])) > smallest
                       and sink["daughters"] > leaf_max_n else SPLIT_SRC)
            if res == SPLIT_SRC and not bs:
                g = node_geom[ent]
                sf = int(g[4])
                if lev >= clevel:
                    appends.append((ANTI_CELL, ent, img))
                    bs = True
                elif sf != 0xFF and lev > 4:
                    # the key cube's integer corner, from its float corner and edge
                    edge = float(g[3])
                    factor = f32(f32(box) / f32(2 ** (lev + 1)))
                    for j in range(8):
                        if sf >> j & 1:
                            continue
                        corner = []
                        for k in range(3):
                            ic = int(round(float(g[k]) / edge)) * 2 + (j >> k & 1)
                            corner.append(f32(factor * f32(ic)))
                        mx = _cube_centre(corner, factor, off)
                        if _outside(mx, cen, cr):
                            appends.append((ANTI_EMPTY + j, ent, img))
                        else:
                            near += 1
            if res != ACCEPT and nb <= leaf_max_n:
                if do_pM and img == central:
                    appends.append((MM, ent, img))
                else:
                    b0 = int(node_info[ent, 0])
                    appends += [(M, -(b + 1), img) for b in range(b0, b0 + nb)]
                res = ACCEPT
                inter += nb
        else:
            appends.append((M, ent, img))
            inter += 1
            res = ACCEPT
            if not bs:
                g = node_geom[int(body_leaf[-ent - 1])]
                mx = _cube_centre(g[:3], g[3], off)
                if _outside(mx, cen, cr):
                    appends.append((ANTI_BODY, ent, img))
                else:
                    near += 1
        results.append(res)
        flags.append(int(bs))
    return results, flags, appends, near, inter


def record_sb(code, ent, img, cells, records, ubs, node_info, bodies, node_geom, body_leaf,
              offsets, uc, rho0):
    """The float32 record an append writes (appendMvec / appendQdvec / appendHdvec,
    nln/mac.c:1421-1505): M-list codes give (mass, x, y, z) at the source's own
    position plus offset; Q / H the cell's record (subtracted moments) with the
    background-flagged variants."""
    off = offsets[img]
    if ent < 0:
        b = bodies[-ent - 1]
        pos = [f32(f32(b[1 + k]) + f32(off[k])) for k in range(3)]
        if code == M:
            return np.array([b[0]] + pos, np.float32)
        assert code == ANTI_BODY
        e = float(node_geom[int(body_leaf[-ent - 1]), 3])
        return np.array([f32(-rho0 * e * e * e)] + pos, np.float32)
    c = cells[ent]
    pos = [f32(f32(c[1 + k]) + f32(off[k])) for k in range(3)]
    lev = int(node_info[ent, 3])
    if code == ANTI_CELL:
        return np.array([-uc[lev][1]] + pos, np.float32)
    if code >= ANTI_EMPTY and code < ANTI_EMPTY + 8:
        return np.array([-uc[lev + 1][1]] + pos, np.float32)
    r = records[ent].copy()
    r[1:4] = pos
    if code == QB:
        r[0] = f32(r[0] + uc[lev][1])
    if code == HB:
        u = ubs[ent]
        r[0], r[20], r[24], r[22] = u[0], u[1], u[2], u[3]
    return r[:13] if code in (Q, QB) else r
