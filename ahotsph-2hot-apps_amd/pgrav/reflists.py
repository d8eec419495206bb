"""Emit a sink group's interaction lists in the reference's own layouts.

The reference's host walk hands the GPU SoA-NSSE blocks (nln/mac.c:117-153:
Hvec / Qvec / Mvec, NSSE = 8 in the AVX build) of absolute-position records,
zero-padded to a multiple of NSSE (nln/mac.c:486-520), and body segments
{base, length} into the body table (MMvec, nln/mac.c:1834-1841).  These
helpers turn the lists of a pgrav.batch.HostPlan group into exactly those
arrays, so the drop-in grav_mn_CUDA / grav_mns_CUDA path can evaluate the same
work as the batched kernel.
"""
import numpy as np

NSSE = 8
SZ = {"hexa": 29, "quad": 13, "mono": 4}


def to_soa(records, nsse=NSSE):
    r = np.asarray(records, np.float32)
    n, sz = r.shape
    nb = (n + nsse - 1) // nsse
    pad = np.zeros((nb * nsse, sz), np.float32)
    pad[:n] = r
    return np.ascontiguousarray(pad.reshape(nb, nsse, sz).transpose(0, 2, 1)).reshape(-1), nb * nsse


def group_soa(H, g, kind):
    """(SoA-8 float32 array, source_n) of a group's `kind` cell list."""
    nodes = H.list(g, kind)
    recs = H.records(nodes)[:, :SZ[kind]] if nodes.size else np.zeros((0, SZ[kind]))
    return to_soa(recs)


def group_segments(H, g):
    """Body list as {base, length} segments over the Morton-sorted body table."""
    b = np.sort(H.list(g, "body").astype(np.int64))
    if b.size == 0:
        return np.zeros((0, 2), np.uint32), 0
    cut = np.flatnonzero(np.diff(b) != 1) + 1
    starts = np.r_[0, cut]
    ends = np.r_[cut, b.size]
    seg = np.stack([b[starts], ends - starts], axis=1).astype(np.uint32)
    return seg, int(b.size)
