"""Restart files (svfsiplus_amd/restart.py) against the record layout the
reference's writer and reader fix (output.cpp:202-376, initialize.cpp:78-241,
the record length of initialize.cpp:503-521): offsets parsed independently
with struct, round trips, multi-rank records, the stamp check and the file
naming, and the bytes and file names of the reference's own write_restart
(compiled from its sources into oracle/_ref) for the same record; a GPU test
continues a run from a restart file bit for bit."""
import os
import struct

import numpy as np
import pytest

from svfsiplus_amd import restart as RS


def _record(tnNo=5, nX=0, dFlag=False, seed=0, cTS=7):
    rng = np.random.default_rng(seed)
    st = RS.make_stamp(tnNo, tDof=4, nX=nX, dFlag=dFlag)
    return RS.Restart(st, cTS, cTS * 1e-3, 12.5, np.array([3.25]), rng.standard_normal((tnNo, 4)),
                      rng.standard_normal((tnNo, 4)), rng.standard_normal(nX),
                      rng.standard_normal((tnNo, 4)) if dFlag else None)


def test_record_length_formula():
    # initialize.cpp:503-521 for the fluid equation: i = 2 tDof (3 tDof with dFlag);
    # recLn = sizeof(int) (1 + 7) + sizeof(double) (2 + nEq + nX + i tnNo)
    assert RS.record_length(RS.make_stamp(5)) == 32 + 8 * (2 + 1 + 0 + 8 * 5)
    assert RS.record_length(RS.make_stamp(5, nX=3, dFlag=True)) == 32 + 8 * (2 + 1 + 3 + 12 * 5)
    assert RS.record_length(RS.make_stamp(209475)) == 32 + 8 * (3 + 8 * 209475)


def test_byte_layout(tmp_path):
    r = _record(tnNo=5, nX=2)
    f = str(tmp_path / "s_007.bin")
    RS.write(f, r)
    b = open(f, "rb").read()
    assert len(b) == RS.record_length(r.stamp)
    assert struct.unpack_from("<7i", b, 0) == (1, 1, 1, 5, 2, 4, 0)   # stamp: ranks nEq nMsh tnNo nX tDof dFlag
    assert struct.unpack_from("<i", b, 28) == (7,)                      # cTS
    assert struct.unpack_from("<2d", b, 32) == (7e-3, 12.5)              # time, CPUT() - timeP(1)
    assert struct.unpack_from("<d", b, 48) == (3.25,)                   # eq.iNorm
    assert struct.unpack_from("<2d", b, 56) == tuple(r.xn)              # cplBC.xn
    # Yn then An, Array(tDof, tnNo) column-major: dof fastest, node by node
    yn = struct.unpack_from("<20d", b, 72)
    an = struct.unpack_from("<20d", b, 72 + 160)
    assert yn == tuple(r.Yn.reshape(-1)) and yn[4:8] == tuple(r.Yn[1])
    assert an == tuple(r.An.reshape(-1))


@pytest.mark.parametrize("nX,dFlag", [(0, False), (3, False), (0, True)])
def test_round_trip(tmp_path, nX, dFlag):
    r = _record(tnNo=11, nX=nX, dFlag=dFlag, seed=nX)
    f = str(tmp_path / "r.bin")
    RS.write(f, r)
    q = RS.read(f, r.stamp)
    assert q.stamp == r.stamp and q.cTS == r.cTS and q.time == r.time and q.cpu_time == r.cpu_time
    for a, b in ((q.iNorm, r.iNorm), (q.xn, r.xn), (q.Yn, r.Yn), (q.An, r.An)):
        assert np.array_equal(a, b)
    assert (q.Dn is None) == (not dFlag)
    if dFlag:
        assert np.array_equal(q.Dn, r.Dn)


def test_rank_records(tmp_path):
    # rank r's record starts at (r - 1) recLn (output.cpp:291-293, initialize.cpp:115-117)
    f = str(tmp_path / "p.bin")
    r1, r2 = _record(seed=1), _record(seed=2)
    RS.write(f, r1, rank=1)
    RS.write(f, r2, rank=2)
    assert os.path.getsize(f) == 2 * RS.record_length(r1.stamp)
    assert np.array_equal(RS.read(f, r1.stamp, rank=2).Yn, r2.Yn)
    assert np.array_equal(RS.read(f, r1.stamp, rank=1).Yn, r1.Yn)


def test_stamp_mismatch(tmp_path):
    r = _record(tnNo=5)
    f = str(tmp_path / "m.bin")
    RS.write(f, r)
    with pytest.raises(RuntimeError, match=r"The number of nodes  5 read from '.*m\.bin' don't match\."):
        RS.read(f, RS.make_stamp(6))
    with pytest.raises(RuntimeError, match="The number of dof "):
        RS.read(f, RS.make_stamp(5, tDof=3))


def test_file_names_and_last_link(tmp_path):
    base = str(tmp_path / "stFile")
    assert RS.file_name(base, 7).endswith("stFile_007.bin")
    assert RS.file_name(base, 1234).endswith("stFile_1234.bin")
    assert RS.file_name(base, 7, replace=True).endswith("stFile_last.bin")
    assert RS.last(base) is None
    f1 = RS.save(base, _record(cTS=1, seed=1))
    f2 = RS.save(base, _record(cTS=2, seed=2))
    assert os.path.basename(f1) == "stFile_001.bin" and os.path.basename(f2) == "stFile_002.bin"
    assert os.path.samefile(RS.last(base), f2)
    assert RS.read(RS.last(base), RS.make_stamp(5)).cTS == 2
    f3 = RS.save(base, _record(cTS=3, seed=3), replace=True)
    assert os.path.basename(f3) == "stFile_last.bin" and RS.read(f3, RS.make_stamp(5)).cTS == 3


def test_shape_checks():
    r = _record()
    r.Yn = r.Yn[:, :3]
    with pytest.raises(ValueError, match="Yn must be"):
        RS.write(os.devnull, r)


@pytest.mark.parametrize("replace,cTS,nX", [(False, 7, 0), (False, 1234, 2), (True, 12, 0)])
def test_bytes_equal_the_reference_writer(tmp_path, replace, cTS, nX):
    """The reference's own write_restart (compiled from its sources into
    oracle/_ref) and restart.save, given the same record: the same file names
    (the numbered file and its _last.bin hard link, or _last.bin alone), the
    same length and the same bytes everywhere but the CPU-time field (bytes
    40-47: utils::cput() at the write in the reference); read back by
    restart.read field for field."""
    import oracle
    if not oracle.ref_available():
        pytest.skip("the reference build (oracle/_ref) is not present")
    r = _record(tnNo=9, nX=nX, cTS=cTS, seed=cTS)
    a, b = tmp_path / "ref", tmp_path / "ours"
    a.mkdir()
    b.mkdir()
    oracle.ref_write_restart(str(a / "stFile"), replace, r.cTS, r.time, r.stamp, r.iNorm, r.xn, r.Yn, r.An,
                             RS.record_length(r.stamp))
    RS.save(str(b / "stFile"), r, replace=replace)
    names = sorted(os.listdir(a))
    assert names == sorted(os.listdir(b))
    assert names == (["stFile_last.bin"] if replace else sorted([os.path.basename(RS.file_name("x", cTS))
                                                                  .replace("x", "stFile"), "stFile_last.bin"]))
    if not replace:
        for d in (a, b):
            assert os.path.samefile(d / names[0], d / names[1])  # a hard link
    ra, rb = (open(d / "stFile_last.bin", "rb").read() for d in (a, b))
    assert len(ra) == len(rb) == RS.record_length(r.stamp)
    assert ra[:40] == rb[:40] and ra[48:] == rb[48:]
    back = RS.read(str(a / "stFile_last.bin"), r.stamp)
    assert back.cTS == r.cTS and back.time == r.time
    assert np.array_equal(back.iNorm, r.iNorm) and np.array_equal(back.xn, r.xn)
    assert np.array_equal(back.Yn, r.Yn) and np.array_equal(back.An, r.An)
