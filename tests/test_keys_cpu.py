"""Body keys (nln/physics_generic.c:77-272): the library's host routines against
the oracle's restatement, and the oracle pinned by the reference's own table,
its two Morton routines agreeing, and the Peano-Hilbert curve's defining
properties."""
import numpy as np
import pytest

from oracle import keys as K


@pytest.fixture(scope="module")
def P():
    from pgrav import keys
    return keys


def test_morton_table_matches_reference_entries():
    # entries of the table printed in nln/physics_generic.c:174-198
    m = K.morton_table()
    listed = {0: 0, 1: 1, 2: 8, 3: 9, 4: 64, 15: 585, 16: 4096, 31: 4681, 32: 32768,
              64: 262144, 127: 299593, 128: 2097152, 200: 2359808, 254: 2396744, 255: 2396745}
    for i, v in listed.items():
        assert m[i] == v, (i, m[i], v)


@pytest.mark.parametrize("exact", [0, 1])
def test_fix_rsize_matches_oracle(P, exact):
    rng = np.random.default_rng(3)
    for _ in range(20):
        lo = rng.normal(size=3).astype(np.float32)
        hi = (lo + rng.uniform(0.1, 3, 3)).astype(np.float32)
        o, rs, kf = P.fix_rsize(lo, hi, exact)
        Rmin, Rsize, keyfactor = K.fix_rsize(lo, hi, exact=bool(exact))
        assert np.array_equal(o, np.float32(Rmin)) and rs == Rsize and kf == keyfactor


def test_keys_match_oracle_and_both_morton_routines_agree(P):
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (400, 3)).astype(np.float32)
    pos[:4] = [[-1, -1, -1], [1, 1, 1], [-1, 1, -1], [0, 0, 0]]  # corners of the box
    lo, hi = pos.min(0), pos.max(0)
    Rmin, Rsize, kf = K.fix_rsize(lo, hi)
    rmin = np.float32(Rmin)
    for kind in (0, 1, 2):
        got = P.keys(pos, rmin, Rsize, kf, kind)
        for p, k in zip(pos, got):
            xp = K.int_pos(p, Rmin, Rsize, kf)
            want = {0: K.get_key, 1: K.get_key_fast, 2: K.get_key_ph}[kind](xp)
            assert k == want, (kind, p, hex(k), hex(want))
            if kind == 0:
                assert want == K.get_key_fast(xp)  # GetKey == GetKeyFast
                assert want >> (3 * K.BITS) == 1  # the leading bit


def test_keys_outside_cube_are_an_error(P):
    from pgrav import PgravError
    Rmin, Rsize, kf = K.fix_rsize([0, 0, 0], [1, 1, 1])
    with pytest.raises(PgravError):
        P.keys(np.array([[0.5, 0.5, 2.0]]), np.float32(Rmin), Rsize, kf, 0)


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_ph_curve_is_a_hilbert_curve(depth):
    """Every cell of a depth once (bijection), consecutive keys face neighbours,
    ph_rev the inverse -- a wrong entry of the automaton breaks one of these."""
    n = 1 << depth
    seen = {}
    for x in range(n):
        for y in range(n):
            for z in range(n):
                k = K.get_key_ph([x, y, z], depth)
                assert k not in seen
                seen[k] = (x, y, z)
                assert K.ph_rev(k) == ([x, y, z], depth)
    keys = sorted(seen)
    assert keys[0] == 1 << (3 * depth) and seen[keys[0]] == (0, 0, 0)
    for a, b in zip(keys, keys[1:]):
        d = np.abs(np.subtract(seen[a], seen[b]))
        assert d.sum() == 1, (seen[a], seen[b])


def test_ph_key_prefix_is_the_parent_cell():
    """A body's PH key at full depth begins with its cell's key at every depth."""
    rng = np.random.default_rng(9)
    for _ in range(50):
        xp = [int(v) for v in rng.integers(0, 1 << K.BITS, 3)]
        full = K.get_key_ph(xp)
        for d in (1, 5, 17):
            assert full >> (3 * (K.BITS - d)) == K.get_key_ph([v >> (K.BITS - d) for v in xp], d)


@pytest.mark.parametrize("ph", [0, 1])
def test_cell_corner_holds_the_body(P, ph):
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 1, (100, 3)).astype(np.float32)
    Rmin, Rsize, kf = K.fix_rsize([0, 0, 0], [1, 1, 1])
    rmin = np.float32(Rmin)
    keys = P.keys(pos, rmin, Rsize, kf, 2 if ph else 0)
    for p, k in zip(pos, keys):
        for d in (1, 4, 9):
            c, sz = P.cell_corner(k >> (3 * (K.BITS - d)), rmin, Rsize, ph)
            edge = Rsize / (1 << d)
            assert np.all(c <= p + 1e-6) and np.all(p < c + edge + 1e-6), (p, c, edge)
            if not ph:
                want, fac = K.cell_corner(int(k >> (3 * (K.BITS - d))), Rmin, Rsize)
                assert np.array_equal(c, np.float32(want)) and sz == fac
            else:
                assert sz == np.float32(Rsize / np.float32(2.0 ** K.BITS))
