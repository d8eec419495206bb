"""SDF snapshots (host code, no GPU): the header grammar of the reference's SDF
writer (sph+nln/wind.c:82-135: '# SDF', the byteorder parameter, one typed
declaration per value, the struct description with its record count, '#\\f',
'# SDF-EOH' padded so the data starts on a 32-byte boundary), the outbody
records of nln/physics.h:49-110 and the header of nln/output.c:270-423."""
import numpy as np
import pytest

from pgrav import sdf
from pgrav._lib import PgravError


def _outbody(n, seed=0, acc=True):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, sdf.OUTBODY_ACC if acc else sdf.OUTBODY)
    for f in r.dtype.names:
        r[f] = np.arange(n) * 3 + 7 if f == "ident" else rng.normal(size=n)
    return r


def test_round_trip(tmp_path):
    recs = _outbody(1001)
    hdr = sdf.output_header(1001, iter_=3, tpos=0.125, tvel=0.0625, tacc=0.1, ke=1.5, pe=-3.25)
    p = tmp_path / "snap.sdf"
    sdf.write(p, recs, sdf.OUTBODY_DESC_ACC, hdr)
    h, got = sdf.read(p)
    assert got.dtype.names == recs.dtype.names and got.dtype.itemsize == 52
    for f in recs.dtype.names:
        np.testing.assert_array_equal(got[f], recs[f])
    assert h["npart"] == 1001 and h["iter"] == 3 and h["tpos"] == 0.125 and h["pe"] == -3.25
    assert h["compiled_version"] == "pgrav"
    assert list(h)[:5] == ["version", "version_2HOT", "units_2HOT", "npart", "iter"]


def test_header_text_and_alignment(tmp_path):
    p = tmp_path / "a.sdf"
    recs = _outbody(5, acc=False)
    sdf.write(p, recs, sdf.OUTBODY_DESC, [("iter", 7), ("npart", (sdf.INT64, 5)),
                                          ("R0", (sdf.FLOAT, 0.5)), ("tpos", 0.1),
                                          ("name", "run a")])
    raw = p.read_bytes()
    text = raw[:raw.index(b"# SDF-EOH")].decode()
    lines = text.split("\n")
    assert lines[0] == "# SDF" and lines[1] == "parameter byteorder = 0x78563412;"
    assert "int iter = 7;" in lines and "int64_t npart = 5;" in lines
    assert "float R0 = 0.5;" in lines            # %g
    assert "double tpos = 0.1000000000000000;" in lines or "double tpos = 0.1;" in lines
    assert 'char name[] = "run a";' in lines
    assert "}[5];" in text and text.endswith("#\f\n")
    start = len(raw) - 5 * sdf.OUTBODY.itemsize
    assert start % 32 == 0 and raw[start - 1:start] == b"\n"
    assert raw[raw.index(b"# SDF-EOH"):start - 1].strip(b" ") == b"# SDF-EOH"


def test_reads_a_header_written_by_hand(tmp_path):
    """A file in the reference writer's format, composed here line by line:
    comments, tabs, an empty record count (taken from the file size), an array
    member."""
    hdr = ("# SDF\nparameter byteorder = 0x78563412;\n# a comment line\n"
           "int iter = 12;\nfloat tolerance = 0.0001;\ndouble tpos = 1.25;\n"
           'char compiled_date[] = "Jan  1 2026";\n'
           "struct {\n\tfloat mass;\t\t/* mass of body */\n\tfloat x, y, z;\n"
           "\tshort flags[2];\n\tint64_t ident;\t/* unique? identifier */\n}[];\n#\f\n# SDF-EOH ")
    pad = (len(hdr) + 1) % 32
    hdr += " " * ((32 - pad) % 32) + "\n"
    dt = np.dtype([("mass", "<f4"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
                   ("flags", "<i2", 2), ("ident", "<i8")])
    recs = np.zeros(3, dt)
    recs["mass"] = [1, 2, 3]
    recs["flags"] = [[1, 2], [3, 4], [5, 6]]
    recs["ident"] = [10, 20, 30]
    p = tmp_path / "h.sdf"
    p.write_bytes(hdr.encode() + recs.tobytes())
    h, got = sdf.read(p)
    assert h == {"iter": 12, "tolerance": pytest.approx(1e-4), "tpos": 1.25,
                 "compiled_date": "Jan  1 2026"}
    assert got.shape == (3,) and got.dtype.itemsize == dt.itemsize
    np.testing.assert_array_equal(got["mass"], [1, 2, 3])
    np.testing.assert_array_equal(got["flags[1]"], [2, 4, 6])
    np.testing.assert_array_equal(got["ident"], [10, 20, 30])


def test_bad_files_are_refused(tmp_path):
    p = tmp_path / "bad.sdf"
    p.write_bytes(b"not an sdf file\n")
    with pytest.raises(PgravError):
        sdf.read(p)
    p.write_bytes(b"# SDF\nparameter byteorder = 0x12345678;\n#\f\n# SDF-EOH \n")
    with pytest.raises(PgravError):
        sdf.read(p)
    p.write_bytes(b"# SDF\nstruct { float m; }[10];\n#\f\n# SDF-EOH \n" + b"\0" * 8)
    with pytest.raises(PgravError):  # declares 10 records, holds 2
        sdf.read(p)
