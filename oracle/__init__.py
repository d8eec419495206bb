"""TEST INFRASTRUCTURE ONLY -- the parity checker.

FP64 CPU restatement of the 2HOT interaction-list arithmetic (pgrav_oracle.c)
plus loaders for the reference's own stand-alone CPU kernels compiled into
oracle/_ref/ (see oracle/Makefile).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package; the product never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double

HSZ, QSZ, MSZ = 29, 13, 4

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_pmn.argtypes = [_i, _i, _f32p, _i, _f32p, _i, _i, _i, _d, _i, _f64p]
        L.orc_pm_sk1.argtypes = [_f32p, _i, _f32p, _i, _i, _d, _f64p]
        L.orc_pmm_sk1.argtypes = [_f32p, _i, _u32p, _i, _f32p, _d, _d, _f64p]
        L.orc_pmm1_mnss_sk1.argtypes = [_f32p, _i, _u16p, _u32p, _i, _u32p, _i, _f32p, _i32p,
                                        _d, _d, _f64p]
        L.orc_pmm1_mnss_sk1.restype = _i
        L.orc_pq1_mnss.argtypes = [_f32p, _i, _u32p, _f32p, _i, _f64p]
        L.orc_pqq1.argtypes = [_f32p, _i, _i32p, _i, _f32p, _i, _f64p]
        L.orc_direct.argtypes = [_f64p, _i64, _f64p, _f64p, _i64, _i, _d, _f64p, _i]
        L.orc_ewald.argtypes = [_f64p, _d, _f64p, ctypes.POINTER(_d)]
        L.orc_direct_ewald.argtypes = [_f64p, _i64, _f64p, _f64p, _i64, _d, _f64p, _i]
        L.orc_ewald_corr.argtypes = [_f64p, _i64, _d, _i, _f64p, _f64p, _f64p, _i]
        L.orc_cell_record.argtypes = [_f64p, _f64p, _i64, _f64p, _d, _f64p]
        L.orc_raw_moments.argtypes = [_f64p, _f64p, _i64, _f64p, _f64p]
        L.orc_mpole.argtypes = [_f64p, _i, _i, _d, _d, _d, _d, _i, _f64p]
        L.orc_eval_records.argtypes = [_f64p, _i64, _f64p, _i64, _i, _i, _d, _i, _f64p]
        L.orc_direct_grad.argtypes = [_f64p, _f64p, _i64, _f64p, _i, _d, _f64p, _i]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _aligned(a, dt=np.float32, align=64):
    """Copy into a 64-byte aligned buffer: the reference SIMD kernels use aligned
    vector loads on their SoA blocks (nln/vec.h vector types)."""
    a = np.ascontiguousarray(a, dtype=dt).reshape(-1)
    raw = np.zeros(a.size * a.itemsize + align, np.uint8)
    off = (-raw.ctypes.data) % align
    out = raw[off:off + a.nbytes].view(dt)
    out[:] = a
    return out


def pmn(order, sink_pos3, f, source_n, sz, nsse=8, dipole=True, eps2=0.0, quad_phi_sign=1):
    """pH/pQ/pM (nln/pgrav.cu:49-410) or do_gravph/pq/p (eps2>0, dipole=False)."""
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    out = np.zeros((m, 4), np.float64)
    lib().orc_pmn(order, int(dipole), s, m, _c(f, np.float32).reshape(-1), source_n, sz, nsse,
                  eps2, quad_phi_sign, out)
    return out


def pm_sk1(sink_pos3, f, source_n, eps_inv, nsse=8):
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    out = np.zeros((m, 4), np.float64)
    lib().orc_pm_sk1(s, m, _c(f, np.float32).reshape(-1), source_n, nsse, eps_inv, out)
    return out


def pmm_sk1(sink_pos3, seg, src_pos3, mmass, eps_inv):
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    seg = _c(seg, np.uint32).reshape(-1)
    out = np.zeros((m, 4), np.float64)
    lib().orc_pmm_sk1(s, m, seg, seg.size // 2, _c(src_pos3, np.float32).reshape(-1), mmass,
                      eps_inv, out)
    return out


def pmm1_mnss_sk1(sink_pos3, ss_index, ss_seg, seg, src_pos3, source_n, mmass, eps_inv):
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    ss_seg = _c(ss_seg, np.uint32).reshape(-1)
    seg = _c(seg, np.uint32).reshape(-1)
    out = np.zeros((m, 4), np.float64)
    bad = lib().orc_pmm1_mnss_sk1(s, m, _c(ss_index, np.uint16), ss_seg, ss_seg.size // 2, seg,
                                  seg.size // 2, _c(src_pos3, np.float32).reshape(-1),
                                  _c(source_n, np.int32), mmass, eps_inv, out)
    return out, bad


def pq1_mnss(sink_pos3, seg, ff, quad_phi_sign=1):
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    out = np.zeros((m, 4), np.float64)
    lib().orc_pq1_mnss(s, m, _c(seg, np.uint32).reshape(-1), _c(ff, np.float32).reshape(-1),
                       quad_phi_sign, out)
    return out


def pqq1(sink_pos3, idx, qtab13, quad_phi_sign=1):
    s = _c(sink_pos3, np.float32).reshape(-1)
    m = s.size // 3
    idx = _c(idx, np.int32)
    out = np.zeros((m, 4), np.float64)
    lib().orc_pqq1(s, m, idx, idx.size, _c(qtab13, np.float32).reshape(-1), quad_phi_sign, out)
    return out


def direct(sink_pos, src_pos, src_mass, soft_type=0, eps=0.0, nthreads=None):
    """FP64 direct sum; soft_type 0 none, 1 Plummer, 4 Dehnen K1."""
    sp = _c(sink_pos, np.float64).reshape(-1)
    q = _c(src_pos, np.float64).reshape(-1)
    mm = _c(src_mass, np.float64).reshape(-1)
    out = np.zeros((sp.size // 3, 4), np.float64)
    lib().orc_direct(sp, sp.size // 3, q, mm, mm.size, soft_type, eps, out,
                     nthreads or min(8, os.cpu_count() or 1))
    return out


def direct_grad(pos, mass, w4, soft_type=1, eps=0.0, nthreads=None):
    """(n, 4): dL/dx, dL/dy, dL/dz, dL/dm of L = sum w.a + v phi (FP64 direct sum)."""
    p = _c(pos, np.float64).reshape(-1)
    m = _c(mass, np.float64).reshape(-1)
    w = _c(w4, np.float64).reshape(-1)
    out = np.zeros((m.size, 4), np.float64)
    lib().orc_direct_grad(p, m, m.size, w, soft_type, eps, out,
                          nthreads or min(8, os.cpu_count() or 1))
    return out


def ewald(x, L=1.0):
    f = np.zeros(3, np.float64)
    phi = _d(0.0)
    lib().orc_ewald(_c(x, np.float64), L, f, ctypes.byref(phi))
    return f, phi.value


def ewald_corr(r, L=1.0, nodes=0, nthreads=None):
    """Ewald correction at r = source - sink (n, 3): val (n, 4) = (C, psi), jac
    (n, 4, 3) with rows dC_x, dC_y, dC_z, dpsi over d(x, y, z), and third (n, 10):
    the second derivatives of C (xxx, xxy, xxz, xyy, xyz, xzz, yyy, yyz, yzz,
    zzz).  nodes = 0: the exact lattice sum at r; nodes > 0: the nearest node of
    the grid of spacing (L/2)/(nodes - 1) and its Taylor steps to r (value to
    third order in psi, second in C, first in the Jacobian; the node's T)."""
    r = _c(r, np.float64).reshape(-1, 3)
    n = r.shape[0]
    val = np.zeros((n, 4), np.float64)
    jac = np.zeros((n, 4, 3), np.float64)
    third = np.zeros((n, 10), np.float64)
    if n:
        lib().orc_ewald_corr(r.reshape(-1), n, L, int(nodes), val.reshape(-1), jac.reshape(-1),
                             third.reshape(-1), nthreads or min(16, os.cpu_count() or 1))
    return val, jac, third


SYM3 = ((0, 0, 0), (0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 1, 2), (0, 2, 2), (1, 1, 1), (1, 1, 2),
        (1, 2, 2), (2, 2, 2))


def sym3_full(third):
    """(n, 10) symmetric components -> (n, 3, 3, 3)."""
    import itertools
    out = np.zeros(third.shape[:-1] + (3, 3, 3))
    for q, (a, b, c) in enumerate(SYM3):
        for p in set(itertools.permutations((a, b, c))):
            out[..., p[0], p[1], p[2]] = third[..., q]
    return out


def direct_ewald(sink_pos, src_pos, src_mass, L=1.0, nthreads=None):
    sp = _c(sink_pos, np.float64).reshape(-1)
    q = _c(src_pos, np.float64).reshape(-1)
    mm = _c(src_mass, np.float64).reshape(-1)
    out = np.zeros((sp.size // 3, 4), np.float64)
    lib().orc_direct_ewald(sp, sp.size // 3, q, mm, mm.size, L, out,
                           nthreads or min(8, os.cpu_count() or 1))
    return out


def _fields(table, names):
    base = table.ctypes.data
    return [ctypes.c_void_p(base + table.dtype.fields[f][1]) for f in names]


def integrate(table, dt, dtv, tpos, tvel, out=None, cosmo=None):
    """nln/integrate.c Integrate (cosmo None) or CosmoIntegrate (cosmo = (a0, a1,
    dt_kick_topos, dt_kick, dt_drift)) on a numpy structured table with fields
    mass, pos, vel, acc, phi; `out` (fields pos, vel) defaults to the table
    itself.  Returns (tpos, tvel, ke, pe)."""
    out = table if out is None else out
    L = lib()
    ke, pe = _d(0.0), _d(0.0)
    fi = _fields(table, ("mass", "pos", "vel", "acc", "phi"))
    fo = _fields(out, ("pos", "vel"))
    if cosmo is None:
        tp, tv = _d(tpos), _d(tvel)
        L.orc_integrate(*fi, table.dtype.itemsize, *fo, out.dtype.itemsize, table.shape[0],
                        _d(dt), _d(dtv), ctypes.byref(tp), ctypes.byref(tv), ctypes.byref(ke),
                        ctypes.byref(pe))
        return tp.value, tv.value, ke.value, pe.value
    L.orc_cosmo_integrate(*fi, table.dtype.itemsize, *fo, out.dtype.itemsize, table.shape[0],
                          *[_d(c) for c in cosmo], ctypes.byref(ke), ctypes.byref(pe))
    return tpos, tvel, ke.value, pe.value


def convert_v(table, scale, to_physical):
    lib().orc_convert_v(*_fields(table, ("vel",)), table.dtype.itemsize, table.shape[0],
                        _d(scale), int(to_physical))


def cell_record(pos, mass, center, bmax):
    rec = np.zeros(HSZ, np.float64)
    p = _c(pos, np.float64).reshape(-1)
    lib().orc_cell_record(p, _c(mass, np.float64), p.size // 3, _c(center, np.float64),
                          float(bmax), rec)
    return rec


def raw_moments(pos, mass, center):
    raw = np.zeros(35, np.float64)
    p = _c(pos, np.float64).reshape(-1)
    lib().orc_raw_moments(p, _c(mass, np.float64), p.size // 3, _c(center, np.float64), raw)
    return raw


def eval_records(sinks, recs, order, dipole=False, eps2=0.0, quad_phi_sign=1):
    """FP64 sum over a list of 29-field records for each sink."""
    s = _c(sinks, np.float64).reshape(-1)
    r = _c(recs, np.float64).reshape(-1)
    out = np.zeros((s.size // 3, 4), np.float64)
    lib().orc_eval_records(s, s.size // 3, r, r.size // 29, order, int(dipole), eps2,
                           quad_phi_sign, out)
    return out


def mpole(rec, order, dx, dipole=True, eps2=0.0, quad_phi_sign=1):
    acc = np.zeros(4, np.float64)
    lib().orc_mpole(_c(rec, np.float64), order, int(dipole), float(dx[0]), float(dx[1]),
                    float(dx[2]), eps2, quad_phi_sign, acc)
    return acc


# ---------------------------------------------------------------- layouts
def to_soa(records, nsse=8):
    """AoS records [n, sz] -> SoA-nsse flat float32 (nln/pgrav.cu:66), zero padded."""
    r = np.asarray(records, np.float32)
    n, sz = r.shape
    nb = (n + nsse - 1) // nsse
    pad = np.zeros((nb * nsse, sz), np.float32)
    pad[:n] = r
    return np.ascontiguousarray(pad.reshape(nb, nsse, sz).transpose(0, 2, 1)).reshape(-1)


# ---------------------------------------------------------------- reference CPU kernels
def ref_available():
    return all(os.path.exists(os.path.join(REF_DIR, n)) for n in
               ("libref_sse4.so", "libref_avx8.so", "libref_vec.so", "libref_pgrav.so"))


_ref = {}


def ref_lib(name):
    """Loads oracle/_ref/libref_<name>.so (built from the reference sources)."""
    if name not in _ref:
        path = os.path.join(REF_DIR, "libref_%s.so" % name)
        if not os.path.exists(path):
            raise RuntimeError("%s not built (make -C oracle ref)" % path)
        _ref[name] = ctypes.CDLL(path)
    return _ref[name]


def ref_mn(kind, sink_body14, source_soa8, nblocks, eps=None):
    """pHinteract / pQinteract / pMinteract / pMinteract_sK1 (nln/pgrav_vec.c) on a
    body table with stride 14 (mass,pos,...,acc at +7..10 relative to &mass)."""
    L = ref_lib("vec")
    p = _c(sink_body14, np.float32).reshape(-1)
    m = p.size // 14
    f = _aligned(source_soa8)
    acc = np.zeros(m * 14 + 8, np.float32)
    fn = getattr(L, {"H": "pHinteract", "Q": "pQinteract", "M": "pMinteract",
                     "MK1": "pMinteract_sK1"}[kind])
    pp = p.ctypes.data_as(ctypes.c_void_p)
    ap = ctypes.c_void_p(acc.ctypes.data + 7 * 4)
    fp = f.ctypes.data_as(ctypes.c_void_p)
    if kind == "MK1":
        ncut = ctypes.c_int(0)
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _i, _i, ctypes.c_void_p, _i,
                       ctypes.c_float, ctypes.POINTER(_i)]
        fn(pp, ap, m, 14, fp, nblocks, ctypes.c_float(eps), ctypes.byref(ncut))
    else:
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, _i, _i, ctypes.c_void_p, _i]
        fn(pp, ap, m, 14, fp, nblocks)
    return acc[:m * 14].reshape(m, 14)[:, 7:11].copy()


def ref_do_grav(name, arch, f_soa, nsources, sink_pos, eps2):
    """do_grav{p,pq,ph,...}_{sse4,avx8} (nln/do_grav_sse4.c, do_grav_avx8.c) for one sink.
    f_soa must be laid out SoA-NSSE with NSSE = 4 (sse4) or 8 (avx8)."""
    L = ref_lib(arch)
    fn = getattr(L, "%s_%s" % (name, arch))
    nsse = 4 if arch == "sse4" else 8
    f = _aligned(f_soa)
    sz = f.size // (((nsources + nsse - 1) // nsse) * nsse) if nsources else 0
    fend = ctypes.c_void_p(f.ctypes.data + 4 * ((nsources + nsse - 1) // nsse) * nsse * sz)
    pos = _c(sink_pos, np.float32)
    mass0 = ctypes.c_float(0)
    acc = np.zeros(3, np.float32)
    phi = ctypes.c_float(0)
    e = ctypes.c_float(eps2)
    ncut = ctypes.c_int(0)
    fn.argtypes = [ctypes.c_void_p] * 8
    fn(f.ctypes.data_as(ctypes.c_void_p), fend, pos.ctypes.data_as(ctypes.c_void_p),
       ctypes.cast(ctypes.byref(mass0), ctypes.c_void_p), acc.ctypes.data_as(ctypes.c_void_p),
       ctypes.cast(ctypes.byref(phi), ctypes.c_void_p),
       ctypes.cast(ctypes.byref(e), ctypes.c_void_p),
       ctypes.cast(ctypes.byref(ncut), ctypes.c_void_p))
    return np.array([acc[0], acc[1], acc[2], phi.value], np.float64)


def ref_pgravm_sk1(sink_p4, source_f4, eps_inv):
    """_pgravm_sK1 (nln/py_pgrav/_pgrav.c): scalar FP64 K1 M x N, stride-4 sinks,
    AoS-4 sources; returns accp (m,4) = (ax, ay, az, phi)."""
    L = ref_lib("pgrav")
    p = _c(sink_p4, np.float32).reshape(-1)
    f = _c(source_f4, np.float32).reshape(-1)
    m, n = p.size // 4, f.size // 4
    acc = np.zeros(m * 4, np.float32)
    ns = ctypes.c_int(0)
    L._pgravm_sK1.argtypes = [_f32p, _f32p, _i, _i, _f32p, _i, ctypes.c_float,
                              ctypes.POINTER(_i)]
    L._pgravm_sK1(p, acc, m, 4, f, n, ctypes.c_float(eps_inv), ctypes.byref(ns))
    return acc.reshape(m, 4)
