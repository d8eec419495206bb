"""Oracle evaluation of the exact lists a plan evaluates (test helper).

For a sink group of a HostPlan, every accepted cell's record is rebuilt by the
oracle from the cell's own bodies (oracle.cell_record, independent of the
library's moment translation) and evaluated in FP64 with oracle.eval_records;
the group's body list goes through the oracle's FP64 direct sum.  Lists of
explicit images (periodic = 2) carry an image per entry (HostPlan.list_images,
nln/mac.c:91-111 numbering): the cell or body is then evaluated at its
position shifted by that image's offset.  The nearest-image periodic mode
(periodic = 1: the Ewald correction per pair) has its own reference,
group_reference_periodic.
"""
import numpy as np

import oracle

ORDER = {"hexa": 4, "quad": 2, "mono": 0}


def image_offsets(images, box, nimage):
    """Source offsets of explicit-image list entries."""
    im = np.asarray(images, np.int64)
    f = 2 * nimage + 1
    off = np.stack([im // (f * f) - nimage, (im // f) % f - nimage, im % f - nimage], -1) * box
    return off.astype(np.float64)


def group_reference(H, g, pos_sorted, mass_sorted, soft="plummer", eps=1e-4, G=1.0,
                    quad_phi_sign=1, images=None):
    """images = (box, nimage) for lists of explicit periodic images."""
    first, cnt, _ = H.group(g)
    slots = H.group_slots(g)  # a run, or a bisection group's body set
    sinks = pos_sorted[slots]
    acc = np.zeros((cnt, 4))
    eps2 = eps * eps if soft == "plummer" else 0.0
    for kind, order in ORDER.items():
        nodes = H.list(g, kind)
        if nodes.size == 0:
            continue
        if images:
            offs = image_offsets(H.list_images(g, kind), *images)
        else:
            offs = np.zeros((nodes.size, 3))
        recs = []
        for nd, o in zip(nodes, offs):
            c = H.cell(int(nd))
            b = pos_sorted[c["first"]:c["first"] + c["count"]] + o
            m = mass_sorted[c["first"]:c["first"] + c["count"]]
            recs.append(oracle.cell_record(b, m, c["cofm"] + o,
                                           c["bmax"] if c["bmax"] > 0 else 1.0))
        acc += oracle.eval_records(sinks, np.array(recs), order, dipole=False, eps2=eps2,
                                   quad_phi_sign=quad_phi_sign)
    bodies = H.list(g, "body")
    if bodies.size:
        if images:
            offs = image_offsets(H.list_images(g, "body"), *images)
        else:
            offs = 0.0
        st = {"none": 0, "plummer": 1, "k1": 4}[soft]
        acc += oracle.direct(sinks, pos_sorted[bodies] + offs, mass_sorted[bodies], soft_type=st,
                             eps=eps)
    return (first if first >= 0 else slots), G * acc


def interactions_cover_all(H, g, n):
    """Sum of body counts over a group's sources == n (nln/mac.c:734)."""
    tot = H.list(g, "body").size
    for kind in ORDER:
        for nd in H.list(g, kind):
            tot += H.cell(int(nd))["count"]
    return tot == n


def _pbc(x, box):
    return x - box * np.rint(x / box)


def group_reference_periodic(H, g, pos_sorted, mass_sorted, box=1.0, soft="none", eps=0.0,
                             G=1.0, nodes=64, ratio=0.05, quad_phi_sign=1, perturb=()):
    """Nearest-image periodic evaluation of group g's own lists (periodic = 1).

    Newtonian part: every cell (multipole record rebuilt from its bodies) and
    body at the image nearest each sink.  Ewald correction (C, psi)(r), r =
    source - sink, from oracle.ewald_corr: the exact lattice sum of nln/ewald.c
    (nodes = 0) or its tabulated form (what the GPU tables hold, nodes = 64):

    * a source whose image is the same for every sink -- |r0_i| < box/2 - group
      radius per axis, r0 = source - group centre (nearest image) -- and whose
      nearest other lattice image is far against the group -- group radius <=
      ratio (box - max_i |r0_i|) -- enters the group's expansion: m (psi, C, J,
      T) at r0 summed over those sources, and
      each sink at offset s from the centre takes
        C(r0 - s) = C - J s + T s s / 2,
        psi(r0 - s) = psi - C.s + s.J.s / 2 - T s s s / 6;
      a cell adds its second moments' coupling at r0: (1/2) T:Q to C and
      (1/2) J:Q to psi, Q its traceless second moment about the centre of mass;
    * the other sources take the correction per pair at r;
    * every cell adds -(2 pi / 3) / box^3 sum m |x - cofm|^2 to the potential
      (the trace part of its second moments against the correction's Hessian).
    A body's own pair has no Newtonian term when unsoftened; the correction
    covers it (at r = 0 its limit C = 0, psi = 2.837297479 / box).

    perturb: planted mistakes, to show what the comparison detects: "scale_c"
    scales C by 1 + 1e-3; "flip_jac" flips the sign of dC_x/dy and dC_y/dx."""
    first, cnt, gc = H.group(g)
    slots = H.group_slots(g)
    sinks = pos_sorted[slots]
    acc = np.zeros((cnt, 4))
    eps2 = eps * eps if soft == "plummer" else 0.0
    lim = 0.5 * box - gc[3]
    rows = np.arange(cnt)

    def expands(r0):
        a = np.max(np.abs(np.atleast_2d(r0)), axis=1)
        return (a < lim) & (gc[3] <= ratio * (box - a))

    pr, pm, prow = [], [], []  # per-pair corrections: r, source mass, sink row
    ex_r0, ex_m, ex_q = [], [], []  # expanded sources: r0, mass, traceless Q (cells)
    for kind, order in ORDER.items():
        for nd in H.list(g, kind):
            c = H.cell(int(nd))
            b = pos_sorted[c["first"]:c["first"] + c["count"]]
            m = mass_sorted[c["first"]:c["first"] + c["count"]]
            rec = oracle.cell_record(b, m, c["cofm"], c["bmax"] if c["bmax"] > 0 else 1.0)
            shift = box * np.rint((sinks - c["cofm"]) / box)
            for sh in np.unique(shift, axis=0):
                sel = np.all(shift == sh, axis=1)
                acc[sel] += oracle.eval_records(sinks[sel] - sh, rec[None], order, dipole=False,
                                                eps2=eps2, quad_phi_sign=quad_phi_sign)
            dlt = b - c["cofm"]
            m2 = np.einsum("n,na,nb->ab", m, dlt, dlt)
            acc[:, 3] += -2.0 * np.pi / 3.0 / box ** 3 * np.trace(m2)
            r0 = _pbc(c["cofm"] - gc[:3], box)
            if expands(r0)[0]:
                ex_r0.append(r0)
                ex_m.append(m.sum())
                ex_q.append(m2 - np.trace(m2) / 3.0 * np.eye(3))
            else:
                pr.append(_pbc(c["cofm"] - sinks, box))
                pm.append(np.full(cnt, m.sum()))
                prow.append(rows)
    bodies = H.list(g, "body")
    if bodies.size:
        bp = pos_sorted[bodies]
        r = _pbc(bp[None, :, :] - sinks[:, None, :], box)  # (cnt, nb, 3)
        r2 = (r * r).sum(-1)
        mb = np.broadcast_to(mass_sorted[bodies][None, :], r2.shape)
        self_pair = r2 == 0.0
        if soft == "plummer":
            s2 = r2 + eps2
        elif soft == "none":
            s2 = np.where(self_pair, 1.0, r2)
        else:
            raise ValueError("periodic reference: soft none or plummer")
        w = np.where(self_pair & (soft == "none"), 0.0, mb)
        rinv = 1.0 / np.sqrt(s2)
        acc[:, :3] += ((w * rinv ** 3)[..., None] * r).sum(1)
        acc[:, 3] -= (w * rinv).sum(1)
        r0 = _pbc(bp - gc[:3], box)
        inside = expands(r0)
        for j in np.nonzero(inside)[0]:
            ex_r0.append(r0[j])
            ex_m.append(mass_sorted[bodies[j]])
            ex_q.append(np.zeros((3, 3)))
        out = ~inside
        if out.any():
            pr.append(r[:, out])
            pm.append(mb[:, out])
            prow.append(np.broadcast_to(rows[:, None], r2.shape)[:, out])

    def planted(val, jac):
        if "scale_c" in perturb:
            val[:, :3] *= 1.0 + 1e-3
        if "flip_jac" in perturb:
            jac[:, 0, 1] *= -1.0
            jac[:, 1, 0] *= -1.0
        return val, jac

    if pr:
        r = np.concatenate([np.reshape(x, (-1, 3)) for x in pr])
        m = np.concatenate([np.reshape(x, -1) for x in pm])
        row = np.concatenate([np.reshape(x, -1) for x in prow])
        val, jac, _ = oracle.ewald_corr(r, L=box, nodes=nodes)
        if "flip_jac" in perturb and nodes:
            # the node step through the flipped Jacobian entries
            h = 0.5 * box / (nodes - 1)
            d = r - np.clip(np.rint(r / h), -(nodes - 1), nodes - 1) * h
            val[:, 0] -= 2.0 * jac[:, 0, 1] * d[:, 1]
            val[:, 1] -= 2.0 * jac[:, 1, 0] * d[:, 0]
        if "scale_c" in perturb:
            val[:, :3] *= 1.0 + 1e-3
        np.add.at(acc, row, m[:, None] * val)
    if ex_r0:
        r0 = np.array(ex_r0)
        m = np.array(ex_m)
        q = np.array(ex_q)
        val, jac, third = oracle.ewald_corr(r0, L=box, nodes=nodes)
        val, jac = planted(val, jac)
        T = oracle.sym3_full(third)
        J = jac[:, :3, :]
        # per source: its mass times the value at r0, plus a cell's second-moment
        # coupling (Q carries the masses already)
        A = (m[:, None] * val[:, :3]).sum(0) + 0.5 * np.einsum("nabc,nbc->a", T, q)
        P = (m * val[:, 3]).sum() + 0.5 * np.einsum("nab,nab->", J, q)
        B = (m[:, None, None] * J).sum(0)
        Tm = (m[:, None, None, None] * T).sum(0)
        s = sinks - gc[:3]
        Ts = np.einsum("abc,nc->nab", Tm, s)
        Tss = np.einsum("nab,nb->na", Ts, s)
        acc[:, :3] += A[None, :] - s @ B.T + 0.5 * Tss
        acc[:, 3] += P - s @ A + 0.5 * np.einsum("na,ab,nb->n", s, B, s) - (Tss * s).sum(1) / 6.0
    return (first if first >= 0 else slots), G * acc
