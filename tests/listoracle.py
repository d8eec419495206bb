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
                             G=1.0, nodes=64, expand=False, quad_phi_sign=1, perturb=()):
    """Nearest-image periodic evaluation of group g's own lists (periodic = 1):
    every cell and body at the image nearest each sink, plus the Ewald
    correction m (C, psi)(r) of each (sink, cell) and (sink, body) pair, r =
    source - sink, from oracle.ewald_corr: the exact lattice sum of nln/ewald.c
    (nodes = 0) or its tabulated form (nearest node of `nodes` per half box plus
    the node Jacobian's first-order step, what the GPU tables hold).  A body's
    own pair has no Newtonian term (unsoftened) and the correction's limit at r =
    0: C = 0 and psi(0) = 2.837297479 / L, its pull from its own images and the
    background.  Each cell's correction is taken at its centre of mass, plus,
    in the potential, the trace part of its second-order term, -(2 pi / 3) /
    box^3 sum m |x - cofm|^2.

    expand: cells whose image is the same for every sink -- |r0_i| < box/2 -
    group radius per axis, r0 = cofm - group centre (nearest image) -- take the
    correction at r0 and its first-order expansion in the sink's offset s from
    the group centre, (C, psi)(r0) - J(r0) s, psi with -(2 pi / 3) |s|^2 /
    box^3 (its second order's trace part); the others and all bodies the
    per-pair form.

    perturb: planted mistakes, to show what the comparison detects:
    "scale_c" scales C by 1 + 1e-3; "flip_jac" flips the sign of dC_x/dy and
    dC_y/dx in every first-order step."""
    first, cnt, gc = H.group(g)
    slots = H.group_slots(g)
    sinks = pos_sorted[slots]
    acc = np.zeros((cnt, 4))
    eps2 = eps * eps if soft == "plummer" else 0.0
    pr, pm, prow = [], [], []  # per-pair corrections: r, source mass, sink row
    er0, em = [], []  # group-centre expansions: r0, cell mass
    lim = 0.5 * box - gc[3]
    rows = np.arange(cnt)
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
            # the correction at the centre of mass leaves, in the potential, the
            # second-order term (1/2) H : sum m d d^T; H's trace is -4 pi / box^3
            acc[:, 3] += -2.0 * np.pi / 3.0 / box ** 3 * (m * ((b - c["cofm"]) ** 2).sum(1)).sum()
            r0 = _pbc(c["cofm"] - gc[:3], box)
            if expand and np.all(np.abs(r0) < lim):
                er0.append(r0)
                em.append(m.sum())
            else:
                pr.append(_pbc(c["cofm"] - sinks, box))
                pm.append(np.full(cnt, m.sum()))
                prow.append(rows)
    bodies = H.list(g, "body")
    if bodies.size:
        r = _pbc(pos_sorted[bodies][None, :, :] - sinks[:, None, :], box)  # (cnt, nb, 3)
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
        pr.append(r)
        pm.append(mb)
        prow.append(np.broadcast_to(rows[:, None], r2.shape))
    h = 0.5 * box / (nodes - 1) if nodes else 0.0

    def planted(val, jac, d):
        if "scale_c" in perturb:
            val[:, :3] *= 1.0 + 1e-3
        if "flip_jac" in perturb and d is not None:
            val[:, 0] -= 2.0 * jac[:, 0, 1] * d[:, 1]
            val[:, 1] -= 2.0 * jac[:, 1, 0] * d[:, 0]
        return val

    if pr:
        r = np.concatenate([np.reshape(x, (-1, 3)) for x in pr])
        m = np.concatenate([np.reshape(x, -1) for x in pm])
        row = np.concatenate([np.reshape(x, -1) for x in prow])
        val, jac = oracle.ewald_corr(r, L=box, nodes=nodes)
        d = r - np.clip(np.rint(r / h), -(nodes - 1), nodes - 1) * h if nodes else None
        val = planted(val, jac, d)
        np.add.at(acc, row, m[:, None] * val)
    if er0:
        r0 = np.array(er0)
        val0, jac0 = oracle.ewald_corr(r0, L=box, nodes=nodes)
        d0 = r0 - np.clip(np.rint(r0 / h), -(nodes - 1), nodes - 1) * h if nodes else None
        val0 = planted(val0, jac0, d0)
        if "flip_jac" in perturb:
            jac0 = jac0.copy()
            jac0[:, 0, 1] *= -1.0
            jac0[:, 1, 0] *= -1.0
        s = sinks - gc[:3]
        s2q = -2.0 * np.pi / 3.0 / box ** 3 * (s * s).sum(1)  # psi's second order, trace part
        for v0, j0, mc in zip(val0, jac0, em):
            acc += mc * (v0[None, :] - s @ j0.T)
            acc[:, 3] += mc * s2q
    return (first if first >= 0 else slots), G * acc
