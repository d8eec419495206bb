"""TEST INFRASTRUCTURE ONLY -- per-call restatement of the reference's DL MAC.

dl_mac(...) answers one DLRcritMAC call (nln/mac.c:1773-1844) for a sink and a
vector of sources: the result per source (0 accept, 1 split sink, 2 split
source) and the records appended, in order, as (list, entity, image) with list
0 = Mvec, 1 = Qvec, 2 = Hvec.  Entities: a tree node (>= 0, a cell) or a body,
-(slot + 1).  Arithmetic in FP32 scalars, as the reference's cell structs and
the MAC's float locals:

  r      = source pos + offset[image]                     (VxVV, :1788)
  dr2    = sum_a (r_a - sink pos_a)^2                     (ddot, :1563)
  accept Mvec  if dr2 > (rcrit + bmax)^2                  (:1796-1799)
         Qvec  elif isquad and dr2 > (rcrit_q + bmax)^2   (:1800-1803)
         Hvec  elif ishexa and dr2 > (rcrit_h + bmax)^2   (:1804-1807)
  else   split sink if dlfac * sink bmax > smallest rcrit of the cell's order,
         else split source                                (:1808-1810)
  a body source: split sink unless the sink is a body, then Mvec   (:1780-1783, :1812-1816)
  a split-source cell: its bodies to Mvec in body order, accepted (:1818-1832);
  with gated, only cells of <= leaf_max_n bodies (the DLRcritMACsb gate, :1740).

isquad / ishexa: p2cut / p4cut nonzero and the cell's body count at least that.
"""
import numpy as np

ACCEPT, SPLIT_SINK, SPLIT_SRC = 0, 1, 2
M, Q, H = 0, 1, 2

f32 = np.float32


def dl_mac(sink, sources, cells, node_info, bodies, offsets, dlfac, p2cut, p4cut, gated,
           leaf_max_n):
    """sink: dict(pos=(3,), bmax, isbody); sources: [(entity, image)];
    cells: (nnodes, 8) float32 mass, x, y, z, rcrit, bmax, rcrit_q, rcrit_h;
    node_info: (nnodes, 4) first body, bodies, has children, level;
    bodies: (n, 4) float32 mass, x, y, z; offsets: (nimage, 3) float32.
    Returns (results, appends, interactions)."""
    sp = [f32(v) for v in sink["pos"]]
    sb = f32(sink["bmax"])
    dl = f32(dlfac)
    results, appends, inter = [], [], 0
    for ent, img in sources:
        ent, img = int(ent), int(img)
        off = offsets[img]
        if ent < 0:
            if not sink["isbody"]:
                results.append(SPLIT_SINK)
                continue
            b = -ent - 1
            appends.append((M, ent, img))
            inter += 1
            results.append(ACCEPT)
            continue
        c = cells[ent]
        mass, rc, bm, rq, rh = (f32(v) for v in (c[0], c[4], c[5], c[6], c[7]))
        r = [f32(f32(c[1 + a]) + f32(off[a])) for a in range(3)]
        d = [f32(r[a] - sp[a]) for a in range(3)]
        dr2 = f32(f32(f32(d[0] * d[0]) + f32(d[1] * d[1])) + f32(d[2] * d[2]))
        nb = int(node_info[ent, 1])
        isquad = p2cut != 0 and nb >= p2cut
        ishexa = p4cut != 0 and nb >= p4cut
        smallest = rh if ishexa else (rq if isquad else rc)
        if dr2 > f32(f32(rc + bm) * f32(rc + bm)):
            appends.append((M, ent, img))
            inter += nb
            res = ACCEPT
        elif isquad and dr2 > f32(f32(rq + bm) * f32(rq + bm)):
            appends.append((Q, ent, img))
            inter += nb
            res = ACCEPT
        elif ishexa and dr2 > f32(f32(rh + bm) * f32(rh + bm)):
            appends.append((H, ent, img))
            inter += nb
            res = ACCEPT
        else:
            res = SPLIT_SINK if f32(dl * sb) > smallest else SPLIT_SRC
        if res == SPLIT_SRC and (not gated or nb <= leaf_max_n):
            res = ACCEPT
            inter += nb
            b0 = int(node_info[ent, 0])
            for b in range(b0, b0 + nb):
                if bodies[b, 0] != bodies[0, 0]:
                    raise ValueError("bad particle mass")
                appends.append((M, -(b + 1), img))
        results.append(res)
    return results, appends, inter


def record(kind, ent, img, cells, records, bodies, offsets):
    """The float32 record an append writes: Mvec (mass, x, y, z); Qvec / Hvec the
    cell's 13 / 29 fields with x, y, z = position + offset (appendQvec / appendHvec,
    nln/mac.c:1506-1561; the dipole, not written there, is zero about the centre of
    mass)."""
    off = offsets[img]
    if ent < 0:
        b = bodies[-ent - 1]
        return np.array([b[0]] + [f32(f32(b[1 + a]) + f32(off[a])) for a in range(3)], np.float32)
    c = cells[ent]
    pos = [f32(f32(c[1 + a]) + f32(off[a])) for a in range(3)]
    if kind == M:
        return np.array([c[0]] + pos, np.float32)
    r = records[ent].copy()
    r[1:4] = pos
    r[5:8] = 0.0
    return r[:13] if kind == Q else r


def soa_records(flat, nrec, sz, nsse=8):
    """(nrec, sz) records from SoA-nsse blocks."""
    a = np.asarray(flat, np.float32).reshape(-1, sz, nsse)
    return a.transpose(0, 2, 1).reshape(-1, sz)[:nrec]
