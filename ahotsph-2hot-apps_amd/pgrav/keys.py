"""Body keys of the reference's tree (include/pgrav.h, host part): the key cube
(nln/physics_generic.c FixRsize, FixRsizeExact), Morton keys (GetKey,
GetKeyFast), Peano-Hilbert keys (GetKeyPH) and cell corners (CellCorner,
CellCornerPH).  Keys come back as Python ints: 1 followed by 31 bits per axis."""
import ctypes

import numpy as np

from ._lib import check, lib, register

_i, _i64, _f = ctypes.c_int, ctypes.c_int64, ctypes.c_float
_fp, _u64p = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_uint64)

register({
    "pgrav_host_fix_rsize": (_i, [_fp, _fp, _i, _fp, _fp, _fp]),
    "pgrav_host_keys": (_i, [_fp, _i64, _fp, _f, _f, _i, _u64p]),
    "pgrav_host_cell_corner": (_i, [_u64p, _i, _fp, _f, _fp, _fp]),
})

MORTON, MORTON_FAST, PH = 0, 1, 2


def _f32(a):
    a = np.ascontiguousarray(a, np.float32)
    return a, a.ctypes.data_as(_fp)


def fix_rsize(rmin, rmax, exact=False):
    """(Rmin float32[3], Rsize, keyfactor) of the key cube around [rmin, rmax]."""
    lo, plo = _f32(rmin)
    hi, phi = _f32(rmax)
    o, po = _f32(np.zeros(3))
    rs, prs = _f32(np.zeros(1))
    kf, pkf = _f32(np.zeros(1))
    lib().pgrav_host_fix_rsize(plo, phi, int(exact), po, prs, pkf)
    check()
    return o, rs[0], kf[0]


def keys(pos, rmin, rsize, keyfactor, kind=MORTON):
    """Key of every body (positions rounded to float, as the reference's body holds them)."""
    p, pp = _f32(np.asarray(pos).reshape(-1, 3))
    r, pr = _f32(rmin)
    out = np.zeros((p.shape[0], 2), np.uint64)
    lib().pgrav_host_keys(pp, p.shape[0], pr, float(rsize), float(keyfactor), int(kind),
                          out.ctypes.data_as(_u64p))
    check()
    return [int(lo) | (int(hi) << 64) for lo, hi in out]


def cell_corner(key, rmin, rsize, ph=False):
    """(corner float32[3], size) of the cell a key names."""
    k = np.array([key & (2 ** 64 - 1), key >> 64], np.uint64)
    r, pr = _f32(rmin)
    c, pc = _f32(np.zeros(3))
    s, ps = _f32(np.zeros(1))
    lib().pgrav_host_cell_corner(k.ctypes.data_as(_u64p), int(ph), pr, float(rsize), pc, ps)
    check()
    return c, s[0]
