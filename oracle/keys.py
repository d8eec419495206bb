"""TEST INFRASTRUCTURE ONLY -- restatement of the reference's body keys.

Python integers and numpy float32 arithmetic, one body at a time, following
nln/physics_generic.c:
  fix_rsize      FixRsize (:77-102) and FixRsizeExact (:104-128)
  int_pos        IntPos (:130-144)
  get_key        GetKey (:146-167): Morton, a leading 1 then x y z bits from the top
  get_key_fast   GetKeyFast (:172-223): the same key from the byte table morton[]
  get_key_ph     GetKeyPH (:226-233) over the automaton setup() builds (:340-451)
                 and phkey (:453-482)
  ph_rev         the inverse of phkey (what phkey_rev, :484-535, is for)
  cell_corner    CellCorner (:236-258)

Pins (tests/test_keys_cpu.py): the byte table against entries the reference's
table lists; get_key == get_key_fast (the reference's two Morton routines);
the PH key is a bijection on every depth's cells, its consecutive cells are
face neighbours (the Hilbert property a wrong transition table breaks) and
ph_rev inverts it; cell_corner of a body's key holds the body.
"""
import numpy as np

BITS = 31  # bits per axis: (KEYBITS - 1) / NDIM with a 96-bit key
RSIZE_EPS = 1e-7
f32 = np.float32


def fix_rsize(rmin, rmax, exact=False):
    """(Rmin[3], Rsize, keyfactor) as float32, statement by statement."""
    rmin = [f32(v) for v in rmin]
    rmax = [f32(v) for v in rmax]
    size = [f32(rmax[i] - rmin[i]) for i in range(3)]
    if exact:
        centre = [f32(f32(rmax[i] + rmin[i]) * f32(0.5)) for i in range(3)]
    else:
        centre = [f32(float(f32(rmax[i] + rmin[i])) * 0.5) for i in range(3)]
    d = max(size[0], size[1], size[2]) if size[0] >= 0 else size[0]
    d = size[0]
    if d < size[1]:
        d = size[1]
    if d < size[2]:
        d = size[2]
    maxchu = f32(2.0 ** BITS)
    if exact:
        Rmin = [f32(f32(f32(-0.5) * d) + centre[i]) for i in range(3)]
        keyfactor = f32(f32(f32(1.0 - RSIZE_EPS) * maxchu) / d)
    else:
        d = f32(float(d) * (1.0 + 2.0 * RSIZE_EPS))
        Rmin = [f32(-0.5 * float(d) + float(centre[i])) for i in range(3)]
        keyfactor = f32(maxchu / d)
    return Rmin, d, keyfactor


def int_pos(p, Rmin, Rsize, keyfactor):
    for a in range(3):
        if f32(p[a]) < Rmin[a] or f32(p[a]) > f32(Rsize + Rmin[a]):
            raise ValueError("position outside the key cube")
    return [int(f32(keyfactor * f32(f32(p[a]) - Rmin[a]))) for a in range(3)]


def get_key(xp):
    key = 1
    for b in range(BITS - 1, -1, -1):
        bits = ((xp[0] >> b) & 1) | (((xp[1] >> b) & 1) << 1) | (((xp[2] >> b) & 1) << 2)
        key = (key << 3) | bits
    return key


def morton_table():
    """morton[i] = i's bits spread three apart (the generator the reference's
    comment gives beside its table)."""
    return [sum(((i >> b) & 1) << (3 * b) for b in range(8)) for i in range(256)]


def get_key_fast(xp):
    m = morton_table()
    k = [m[(xp[0] >> (8 * b)) & 0xff] | m[(xp[1] >> (8 * b)) & 0xff] << 1 |
         m[(xp[2] >> (8 * b)) & 0xff] << 2 for b in range(4)]
    lo = (((k[2] & 0xffff) << 48) | (k[1] << 24) | k[0]) & (2 ** 64 - 1)
    hi = (1 << 29) | ((k[3] << 8) & 0xffffffff) | (k[2] >> 16)
    return lo | (hi << 64)


# the automaton: start corners with an even number of bits set, three path types
_S_MASK = [0, 3, 5, 6]
_MASK_S = {0: 0, 3: 1, 5: 2, 6: 3}
_P_MASK = [4, 2, 1]
_BIG = [[1, 2, 1, 0, 1, 2, 1, 0], [0, 2, 0, 1, 0, 2, 0, 1], [0, 1, 0, 2, 0, 1, 0, 2]]
_LTL = [[1, 2, 2, 0, 0, 2, 2, 1], [0, 2, 2, 1, 1, 2, 2, 0], [0, 1, 1, 2, 2, 1, 1, 0]]


def ph_tables():
    bitmap, revbitmap, startmap, typmap = {}, {}, {}, {}
    for typ in range(3):
        for start in range(4):
            bm = st = _S_MASK[start]
            smap, bmap, lmap = [0] * 8, [0] * 8, [0] * 8
            for j in range(8):
                smap[j] = _MASK_S[st]
                lmap[j] = _LTL[typ][j]
                bmap[j] = bm
                bt = _P_MASK[_BIG[typ][j]]
                lt = _P_MASK[_LTL[typ][j]]
                bm ^= bt
                st ^= bt ^ lt
            for j in range(8):
                b = bmap[j]
                startmap[start, typ, b] = smap[j]
                bitmap[start, typ, b] = j
                revbitmap[start, typ, j] = b
                typmap[start, typ, b] = lmap[j]
    return bitmap, revbitmap, startmap, typmap


def get_key_ph(xp, depth=BITS):
    bitmap, _, startmap, typmap = ph_tables()
    key, start, typ = 1, 0, 0
    for b in range(depth - 1, -1, -1):
        bits = ((xp[0] >> b) & 1) | (((xp[1] >> b) & 1) << 1) | (((xp[2] >> b) & 1) << 2)
        key = (key << 3) | bitmap[start, typ, bits]
        start, typ = startmap[start, typ, bits], typmap[start, typ, bits]
    return key


def ph_rev(key):
    """Integer corner (in units of the cell at the key's depth) and the depth."""
    _, revbitmap, startmap, typmap = ph_tables()
    depth = 0
    while key >> (3 * (depth + 1)):
        depth += 1
    out, start, typ = [0, 0, 0], 0, 0
    for d in range(depth - 1, -1, -1):
        j = (key >> (3 * d)) & 7
        o = revbitmap[start, typ, j]
        for a in range(3):
            out[a] |= ((o >> a) & 1) << d
        start, typ = startmap[start, typ, o], typmap[start, typ, o]
    return out, depth


def cell_corner(key, Rmin, Rsize):
    icorner, iscale = [0, 0, 0], 1
    while key > 1:
        for a in range(3):
            if key & (1 << a):
                icorner[a] |= iscale
        key >>= 3
        iscale <<= 1
    factor = f32(Rsize / f32(iscale))
    return [f32(Rmin[a] + f32(factor * f32(icorner[a]))) for a in range(3)], factor
