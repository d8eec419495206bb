"""SDF snapshots of the body table (libpgrav part 5).

The outbody records and header of nln/output.c (SDFwrite64 call :270-423,
OUTBODYDESC of nln/physics.h:59-110) written and read through the library's
host SDF code; records come back as numpy structured arrays whose fields are
the struct members the file declares.
"""
import ctypes

import numpy as np

from ._lib import PgravError, check, lib, register

INT, INT64, FLOAT, DOUBLE, STRING, CHAR, SHORT = range(7)
_NP = {INT: "<i4", INT64: "<i8", FLOAT: "<f4", DOUBLE: "<f8", CHAR: "i1", SHORT: "<i2"}


class Value(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("type", ctypes.c_int), ("i", ctypes.c_int64),
                ("d", ctypes.c_double), ("s", ctypes.c_char_p)]


_vp, _i, _i64, _cp = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_char_p
register({
    "pgrav_sdf_write": (_i, [_cp, ctypes.POINTER(Value), _i, _cp, _vp, _i64, _i64]),
    "pgrav_sdf_open": (_vp, [_cp]),
    "pgrav_sdf_close": (None, [_vp]),
    "pgrav_sdf_nvalues": (_i, [_vp]),
    "pgrav_sdf_value_at": (_i, [_vp, _i, ctypes.POINTER(Value)]),
    "pgrav_sdf_nrecords": (_i64, [_vp]),
    "pgrav_sdf_recsize": (_i64, [_vp]),
    "pgrav_sdf_nfields": (_i, [_vp]),
    "pgrav_sdf_field_at": (_i, [_vp, _i, ctypes.POINTER(_cp), ctypes.POINTER(_i),
                                ctypes.POINTER(_i64)]),
    "pgrav_sdf_read": (_i, [_vp, _i64, _i64, _vp]),
})

# nln/physics.h:59-68 (SAVE_ACC) and :86-92, the text written into the header
OUTBODY_DESC_ACC = """struct {
    float mass;\t\t\t/* mass of body */
    float x, y, z;\t\t/* position of body */
    float vx, vy, vz;\t\t/* velocity of body */
    float ax, ay, az;\t\t/* acceleration */
    float phi;\t\t\t/* potential */
    int64_t ident;\t\t/* unique? identifier */
}"""
OUTBODY_DESC = """struct {
    float mass;\t\t\t/* mass of body */
    float x, y, z;\t\t/* position of body */
    float vx, vy, vz;\t\t/* velocity of body */
    int64_t ident;\t\t/* unique? identifier */
}"""
OUTBODY_ACC = np.dtype([("mass", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                        ("vx", "<f4"), ("vy", "<f4"), ("vz", "<f4"), ("ax", "<f4"), ("ay", "<f4"),
                        ("az", "<f4"), ("phi", "<f4"), ("ident", "<i8")])
OUTBODY = np.dtype([("mass", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("vx", "<f4"),
                    ("vy", "<f4"), ("vz", "<f4"), ("ident", "<i8")])


def _value(name, v):
    """(name, value) or (name, type, value) -> Value; the type follows the Python
    value (int: int, or int64 beyond 32 bits; float: double; str: string)."""
    if isinstance(v, tuple):
        t, v = v
    elif isinstance(v, (bool, int, np.integer)):
        t = INT if -2 ** 31 <= int(v) < 2 ** 31 else INT64
    elif isinstance(v, (float, np.floating)):
        t = DOUBLE
    elif isinstance(v, str):
        t = STRING
    else:
        raise TypeError("header value %s: %r" % (name, v))
    out = Value(name.encode(), t, 0, 0.0, None)
    if t in (INT, INT64):
        out.i = int(v)
    elif t in (FLOAT, DOUBLE):
        out.d = float(v)
    else:
        out.s = str(v).encode()
    return out


def write(path, records, desc, header=()):
    """Writes `records` (a packed numpy structured array) under the struct text
    `desc`; header: (name, value) or (name, (type, value)) pairs, in order."""
    records = np.ascontiguousarray(records)
    vals = (Value * max(1, len(header)))(*[_value(n, v) for n, v in header])
    rc = lib().pgrav_sdf_write(str(path).encode(), vals, len(header), desc.encode(),
                               records.ctypes.data if records.size else None, records.shape[0],
                               records.dtype.itemsize)
    check()
    if rc:
        raise PgravError("pgrav_sdf_write failed")


def read(path):
    """(header dict, records) of an SDF file."""
    L = lib()
    h = L.pgrav_sdf_open(str(path).encode())
    check()
    if not h:
        raise PgravError("pgrav_sdf_open failed")
    try:
        header = {}
        v = Value()
        for k in range(L.pgrav_sdf_nvalues(h)):
            L.pgrav_sdf_value_at(h, k, ctypes.byref(v))
            name = v.name.decode()
            header[name] = (v.i if v.type in (INT, INT64) else
                            v.d if v.type in (FLOAT, DOUBLE) else v.s.decode())
        names, formats, offsets = [], [], []
        nm, ty, off = _cp(), _i(), _i64()
        for k in range(L.pgrav_sdf_nfields(h)):
            L.pgrav_sdf_field_at(h, k, ctypes.byref(nm), ctypes.byref(ty), ctypes.byref(off))
            names.append(nm.value.decode())
            formats.append(_NP[ty.value])
            offsets.append(off.value)
        n = L.pgrav_sdf_nrecords(h)
        dtype = np.dtype({"names": names, "formats": formats, "offsets": offsets,
                          "itemsize": max(1, L.pgrav_sdf_recsize(h))})
        recs = np.zeros(n, dtype)
        if n:
            rc = L.pgrav_sdf_read(h, 0, n, recs.ctypes.data)
            check()
            if rc:
                raise PgravError("pgrav_sdf_read failed")
        return header, recs
    finally:
        L.pgrav_sdf_close(h)


def output_header(npart, iter_=0, tpos=0.0, tvel=0.0, tacc=0.0, R=(0.5, 0.5, 0.5), N=(0, 0, 0),
                  do_periodic=0, eps=0.0, eps_scaled=0.0, smoothing=1, tol=0.0, ke=0.0, pe=0.0,
                  G=1.0):
    """The header nln/output.c:270-423 writes for a run without cosmology (a, z,
    H = 0 as :226-229 sets them; the cosmology's parameters 0, G = Gnewt)."""
    zero = 0.0
    return [("version", 2), ("version_2HOT", 2), ("units_2HOT", 2), ("npart", (INT64, npart)),
            ("iter", iter_), ("do_periodic", do_periodic), ("tpos", tpos), ("tvel", tvel),
            ("tacc", tacc), ("R0", (FLOAT, R[2])), ("L0", (FLOAT, 2.0 * R[2])),
            ("redshift", zero), ("a", zero), ("a_tvel", zero), ("a_tacc", zero),
            ("Omega0", zero), ("Omega0_m", zero), ("Omega0_r", zero), ("Omega0_lambda", zero),
            ("Omega0_cdm", zero), ("Omega0_ncdm_tot", zero), ("Omega0_b", zero),
            ("Omega0_g", zero), ("Omega0_ur", zero), ("Omega0_fld", zero), ("w0_fld", zero),
            ("wa_fld", zero), ("h_100", zero), ("H0", zero), ("hubble", zero), ("H", zero),
            ("Gnewt", float(G)), ("growthfac", zero), ("growthfac_tvel", zero),
            ("growthfac_tacc", zero), ("velfac", zero), ("tolerance", (FLOAT, tol)),
            ("frac_tolerance", (FLOAT, 0.0)), ("frac_tolerance0", (FLOAT, 0.0)),
            ("Rx", (FLOAT, R[0])), ("Ry", (FLOAT, R[1])), ("Rz", (FLOAT, R[2])),
            ("Nx", N[0]), ("Ny", N[1]), ("Nz", N[2]), ("epsilon_mscale", (FLOAT, eps)),
            ("epsilon_scaled", (FLOAT, eps_scaled)), ("epsilon0", (FLOAT, 0.0)),
            ("force_smoothing_type", smoothing), ("ic_Nmesh", 0), ("ic_growthfac", zero),
            ("checkpoint", 0), ("npart_orig", (INT64, npart)), ("subsample_fraction", zero),
            ("subsample_random_seed", 0), ("ke", float(ke)), ("pe", float(pe)),
            ("compiled_version", "pgrav"), ("compiled_date", ""), ("compiled_time", "")]
