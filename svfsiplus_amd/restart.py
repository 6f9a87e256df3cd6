"""svFSIplus restart files (`.bin`) for the fluid equation.

The reference writes one fixed-length record per rank (output.cpp:202-360,
write_restart, with the header of write_restart_header, output.cpp:365-376)
and reads it back at start-up (initialize.cpp:78-241, init_from_bin, with
read_restart_header, output.cpp:182-195). The record of rank r (1-based)
starts at byte (r - 1) * recLn, native byte order (little-endian here):

    stamp      7 x int32   {ranks, nEq, nMsh, tnNo, nX, tDof, dFlag}  (initialize.cpp:497)
    cTS        int32       time step number
    time       float64     physical time
    cpu_time   float64     CPUT() - timeP(1) at the write
    iNorm      nEq x float64
    xn         nX x float64     coupled-BC state (cplBC.xn)
    Yn         tDof x tnNo float64, dof fastest (Array(tDof, tnNo))
    An         tDof x tnNo float64
    Dn         tDof x tnNo float64, only with dFlag (a moving mesh)

recLn = 4 (1 + 7) + 8 (2 + nEq + nX + k tDof tnNo), k = 2 (3 with dFlag)
(initialize.cpp:503-521; the fluid equation has no prestress, shell, CEP or
immersed-body fields). Reading fills the *old* state (Yo, Ao; Do), so the run
continues with time step cTS + 1, and compares the stamp with the run's own;
a mismatch is an error naming the first field that differs, as
initialize.cpp:205-230 reports it.

File names follow output.cpp:271-279: `<stFileName>_<cTS>.bin` (cTS as %03d
below 1000) with a hard link `<stFileName>_last.bin` to the newest, or
`<stFileName>_last.bin` alone when every save replaces the last one; a run
that continues a previous simulation starts from `<stFileName>_last.bin`
(initialize.cpp:655-662).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

# initialize.cpp:205-213: the field names of the stamp in its mismatch message
STAMP_NAMES = ("The number of processors ", "The number of equations ", "The number of meshes ",
               "The number of nodes ", "The number of ncplBC.x ", "The number of dof ",
               "The dFlag specification ")


def make_stamp(tnNo: int, tDof: int = 4, nEq: int = 1, nMsh: int = 1, nX: int = 0, nprocs: int = 1,
               dFlag: bool = False) -> tuple:
    """The run's stamp (initialize.cpp:487-497)."""
    return (int(nprocs), int(nEq), int(nMsh), int(tnNo), int(nX), int(tDof), 1 if dFlag else 0)


def record_length(stamp: tuple) -> int:
    """Bytes of one rank's record (initialize.cpp:503-521) for the fluid equation."""
    _, nEq, _, tnNo, nX, tDof, dFlag = stamp
    k = (3 if dFlag else 2) * tDof
    return 4 * (1 + len(stamp)) + 8 * (2 + nEq + nX + k * tnNo)


def file_name(stFileName: str, cTS: int, replace: bool = False) -> str:
    """output.cpp:262-279."""
    if replace:
        return stFileName + "_last.bin"
    return stFileName + "_" + (f"{cTS}" if cTS >= 1000 else f"{cTS:03d}") + ".bin"


@dataclass
class Restart:
    """One rank's record. Yn/An/Dn are (tnNo, tDof): node-major with the dof
    fastest, the memory order of the reference's Array(tDof, tnNo)."""
    stamp: tuple
    cTS: int
    time: float
    cpu_time: float
    iNorm: np.ndarray
    Yn: np.ndarray
    An: np.ndarray
    xn: np.ndarray = field(default_factory=lambda: np.zeros(0))
    Dn: np.ndarray | None = None

    def check(self):
        nprocs, nEq, nMsh, tnNo, nX, tDof, dFlag = self.stamp
        if len(self.stamp) != 7:
            raise ValueError("restart: the stamp has 7 fields")
        if np.asarray(self.iNorm).shape != (nEq,):
            raise ValueError(f"restart: iNorm must hold nEq = {nEq} values")
        if np.asarray(self.xn).shape != (nX,):
            raise ValueError(f"restart: xn must hold nX = {nX} values")
        for name, a in (("Yn", self.Yn), ("An", self.An)) + ((("Dn", self.Dn),) if dFlag else ()):
            if a is None or np.asarray(a).shape != (tnNo, tDof):
                raise ValueError(f"restart: {name} must be ({tnNo}, {tDof})")


def _record(r: Restart) -> bytes:
    r.check()
    parts = [np.asarray(r.stamp, dtype="<i4").tobytes(), np.asarray([r.cTS], dtype="<i4").tobytes(),
             np.asarray([r.time, r.cpu_time], dtype="<f8").tobytes(),
             np.asarray(r.iNorm, dtype="<f8").tobytes(), np.asarray(r.xn, dtype="<f8").tobytes(),
             np.ascontiguousarray(r.Yn, dtype="<f8").tobytes(), np.ascontiguousarray(r.An, dtype="<f8").tobytes()]
    if r.stamp[6]:
        parts.append(np.ascontiguousarray(r.Dn, dtype="<f8").tobytes())
    b = b"".join(parts)
    assert len(b) == record_length(r.stamp)
    return b


def write(fname: str, r: Restart, rank: int = 1):
    """Write rank `rank`'s record (1-based) at (rank - 1) * recLn, creating the
    file if needed (output.cpp:281-300)."""
    rec = _record(r)
    mode = "r+b" if os.path.exists(fname) and rank > 1 else "wb"
    with open(fname, mode) as f:
        f.seek((rank - 1) * len(rec))
        f.write(rec)


def read(fname: str, stamp: tuple, rank: int = 1) -> Restart:
    """Read rank `rank`'s record of a run whose own stamp is `stamp`; a file
    written with another stamp is refused (initialize.cpp:205-230)."""
    nprocs, nEq, nMsh, tnNo, nX, tDof, dFlag = stamp
    recLn = record_length(stamp)
    with open(fname, "rb") as f:
        f.seek((rank - 1) * recLn)
        buf = f.read(recLn)
    if len(buf) < 4 * 7:
        raise RuntimeError(f"restart: '{fname}' is shorter than a record header")
    t = tuple(int(v) for v in np.frombuffer(buf, dtype="<i4", count=7))
    for i, (a, b) in enumerate(zip(t, stamp)):
        if a != b:
            raise RuntimeError(f"{STAMP_NAMES[i]} {a} read from '{fname}' don't match.")
    if len(buf) < recLn:
        raise RuntimeError(f"restart: '{fname}' holds {len(buf)} of the record's {recLn} bytes")
    o = 28
    cTS = int(np.frombuffer(buf, dtype="<i4", count=1, offset=o)[0])
    o += 4
    time, cpu = (float(v) for v in np.frombuffer(buf, dtype="<f8", count=2, offset=o))
    o += 16

    def take(count, shape=None):
        nonlocal o
        a = np.frombuffer(buf, dtype="<f8", count=count, offset=o).copy()
        o += 8 * count
        return a.reshape(shape) if shape else a

    iNorm = take(nEq)
    xn = take(nX)
    Yn = take(tnNo * tDof, (tnNo, tDof))
    An = take(tnNo * tDof, (tnNo, tDof))
    Dn = take(tnNo * tDof, (tnNo, tDof)) if dFlag else None
    return Restart(t, cTS, time, cpu, iNorm, Yn, An, xn, Dn)


def save(stFileName: str, r: Restart, replace: bool = False) -> str:
    """write_restart's file handling (output.cpp:262-279, 353-358): the
    numbered file plus a hard link `_last.bin` to it, or `_last.bin` alone."""
    fname = file_name(stFileName, r.cTS, replace)
    write(fname, r)
    if not replace:
        last = stFileName + "_last.bin"
        if os.path.lexists(last):
            os.remove(last)
        os.link(fname, last)
    return fname


def last(stFileName: str) -> str | None:
    """The file a continued simulation starts from (initialize.cpp:655-662), or
    None when there is none (the reference then starts from zero)."""
    f = stFileName + "_last.bin"
    return f if os.path.exists(f) else None
