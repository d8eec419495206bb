"""Oracle evaluation of the exact lists a plan evaluates (test helper).

For a sink group of a HostPlan, every accepted cell's record is rebuilt by the
oracle from the cell's own bodies (oracle.cell_record, independent of the
library's moment translation) and evaluated in FP64 with oracle.eval_records;
the group's body list goes through the oracle's FP64 direct sum.  The periodic
Ewald correction is not included here (see test_periodic_gpu.py).  Lists of
explicit images (periodic = 2) carry an image per entry (HostPlan.list_images,
nln/mac.c:91-111 numbering): the cell or body is then evaluated at its
position shifted by that image's offset.
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
