"""The step either side of the force evaluation: the leapfrog kick-drift of the
body table on the GPU (libpgrav part 4; nln/integrate.c).

Tables are numpy structured arrays (host) or device buffers laid out the same
way: float fields with one byte stride, the reference's packed `body`
(nln/physics.h:30-44) by default.
"""
import ctypes

import numpy as np

from ._lib import PgravError, check, lib, register

# nln/physics.h:30-44, packed, without the optional key
BODY = np.dtype([("mass", "<f4"), ("pos", "<f4", 3), ("vel", "<f4", 3), ("acc", "<f4", 3),
                 ("phi", "<f4"), ("nterms", "<f4"), ("ident", "<i8")])

_vp, _i, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
register({
    "pgrav_integrate": (_i, [_vp] * 5 + [_i, _vp, _vp, _i, _i, _d, _d, _dp, _dp, _dp, _dp]),
    "pgrav_integrate_device": (_i, [_vp] * 5 + [_i, _vp, _vp, _i, _i, _d, _d, _dp, _dp, _dp, _dp,
                                                _vp]),
    "pgrav_integrate_device_noalloc": (_i, [_vp] * 5 + [_i, _vp, _vp, _i, _i, _d, _d, _dp, _dp,
                                                        _dp, _dp, _vp]),
    "pgrav_cosmo_integrate_device": (_i, [_vp] * 5 + [_i, _vp, _vp, _i, _i] + [_d] * 5 +
                                     [_dp, _dp, _vp]),
    "pgrav_convert_v_device": (_i, [_vp, _i, _i, _d, _i, _vp]),
})

IN_FIELDS = ("mass", "pos", "vel", "acc", "phi")
OUT_FIELDS = ("pos", "vel")


def _ptrs(base, dtype, names):
    return [base + dtype.fields[f][1] for f in names]


def _done(rc, what):
    check()
    if rc:
        raise PgravError(what + " failed")


def integrate(table, dt, dtv, tpos, tvel, out=None):
    """Integrate (nln/integrate.c:16-59) on host tables, through the GPU.
    Returns (tpos, tvel, ke, pe); `out` (fields pos, vel) defaults to `table`."""
    out = table if out is None else out
    tp, tv, ke, pe = _d(tpos), _d(tvel), _d(), _d()
    rc = lib().pgrav_integrate(*_ptrs(table.ctypes.data, table.dtype, IN_FIELDS),
                               table.dtype.itemsize,
                               *_ptrs(out.ctypes.data, out.dtype, OUT_FIELDS), out.dtype.itemsize,
                               table.shape[0], dt, dtv, ctypes.byref(tp), ctypes.byref(tv),
                               ctypes.byref(ke), ctypes.byref(pe))
    _done(rc, "pgrav_integrate")
    return tp.value, tv.value, ke.value, pe.value


def integrate_device(ptr, n, dtype, dt, dtv, tpos, tvel, out_ptr=None, out_dtype=None, stream=None,
                     noalloc=False):
    """Integrate on a device table at `ptr` (n records of `dtype`).  noalloc:
    through pgrav_integrate_device_noalloc (no device allocation; n <= 2^27)."""
    out_ptr = ptr if out_ptr is None else out_ptr
    out_dtype = dtype if out_dtype is None else out_dtype
    tp, tv, ke, pe = _d(tpos), _d(tvel), _d(), _d()
    name = "pgrav_integrate_device_noalloc" if noalloc else "pgrav_integrate_device"
    rc = getattr(lib(), name)(*_ptrs(ptr, dtype, IN_FIELDS), dtype.itemsize,
                              *_ptrs(out_ptr, out_dtype, OUT_FIELDS), out_dtype.itemsize,
                              n, dt, dtv, ctypes.byref(tp), ctypes.byref(tv),
                              ctypes.byref(ke), ctypes.byref(pe), stream)
    _done(rc, name)
    return tp.value, tv.value, ke.value, pe.value


def cosmo_integrate_device(ptr, n, dtype, a0, a1, dt_kick_topos, dt_kick, dt_drift, out_ptr=None,
                           out_dtype=None, stream=None):
    """CosmoIntegrate (nln/integrate.c:82-135) with the cosmology's factors given.
    Returns (ke, pe)."""
    out_ptr = ptr if out_ptr is None else out_ptr
    out_dtype = dtype if out_dtype is None else out_dtype
    ke, pe = _d(), _d()
    rc = lib().pgrav_cosmo_integrate_device(*_ptrs(ptr, dtype, IN_FIELDS), dtype.itemsize,
                                            *_ptrs(out_ptr, out_dtype, OUT_FIELDS),
                                            out_dtype.itemsize, n, a0, a1, dt_kick_topos, dt_kick,
                                            dt_drift, ctypes.byref(ke), ctypes.byref(pe), stream)
    _done(rc, "pgrav_cosmo_integrate_device")
    return ke.value, pe.value


def convert_v_device(ptr, n, dtype, scale, to_physical, stream=None):
    """ConvertV (nln/integrate.c:62-80) on a device table."""
    rc = lib().pgrav_convert_v_device(ptr + dtype.fields["vel"][1], dtype.itemsize, n, scale,
                                      int(to_physical), stream)
    _done(rc, "pgrav_convert_v_device")
