"""VTK XML meshes and results for the fluid path: the .vtu volume mesh and the
.vtp face files a SimVascular model directory holds, and .vtu result files.

The reference reads them through VTK's XML readers (vtk_xml_parser.cpp:
load_vtu, load_vtp) and writes results with vtk_xml.cpp (write_vtus). This
module reads and writes the same files with numpy, zlib and base64 only:
  - every DataArray encoding VTK writes: ascii, inline base64 ("binary") and
    appended data (raw or base64), each with or without the zlib block
    compression of vtkZLibDataCompressor, UInt32 or UInt64 headers;
  - the conventions of load_vtp (vtk_xml_parser.cpp:466-620): a face's
    GlobalNodeID and GlobalElementID are 1-based ids into the volume mesh
    (gN = GlobalNodeID - 1, gE = GlobalElementID - 1).

    mesh = load_mesh("mesh-complete.mesh.vtu",
                     {"inflow": "mesh-surfaces/inflow.vtp", ...},
                     dirichlet=["inflow", "walls"])
    write_result("result_001.vtu", mesh, Yn, An)
"""
from __future__ import annotations

import base64
import os
import xml.etree.ElementTree as ET
import zlib
from dataclasses import dataclass, field

import numpy as np

VTK_TRIANGLE, VTK_TETRA = 5, 10
_TYPES = {"Int8": np.int8, "UInt8": np.uint8, "Int16": np.int16, "UInt16": np.uint16, "Int32": np.int32,
          "UInt32": np.uint32, "Int64": np.int64, "UInt64": np.uint64, "Float32": np.float32, "Float64": np.float64}
_NAMES = {np.dtype(v): k for k, v in _TYPES.items()}


@dataclass
class VtkData:
    """One piece of an UnstructuredGrid (.vtu) or PolyData (.vtp) file."""
    kind: str                          # "UnstructuredGrid" or "PolyData"
    points: np.ndarray                 # (nNo, 3) float64
    connectivity: np.ndarray           # flat point ids of every cell
    offsets: np.ndarray                # end offset of each cell in connectivity
    types: np.ndarray | None           # VTK cell types (UnstructuredGrid); None for PolyData polys
    point_data: dict = field(default_factory=dict)
    cell_data: dict = field(default_factory=dict)

    def cells(self, size: int) -> np.ndarray:
        """The cells as an (nCells, size) array (every cell must have `size` points)."""
        n = np.diff(np.concatenate([[0], self.offsets]))
        if not np.all(n == size):
            raise ValueError(f"cells of {sorted(set(n.tolist()))} points, expected {size}")
        return self.connectivity.reshape(-1, size)


# ------------------------------------------------------------------ reading --
def _b64(text: bytes, start: int, nbytes: int) -> tuple[bytes, int]:
    """nbytes decoded from the base64 text at `start` and the position after them."""
    nchar = 4 * ((nbytes + 2) // 3)
    return base64.b64decode(text[start:start + nchar]), start + nchar


def _block(buf: bytes | None, text: bytes | None, pos: int, hs: int, compressed: bool) -> bytes:
    """The payload of one DataArray: a header (byte count, or the block table
    when compressed) then the data, from raw bytes (buf) or base64 text."""
    hdt = np.uint32 if hs == 4 else np.uint64
    if buf is not None:  # raw
        if not compressed:
            n = int(np.frombuffer(buf, hdt, 1, pos)[0])
            return buf[pos + hs:pos + hs + n]
        nb = int(np.frombuffer(buf, hdt, 1, pos)[0])
        head = np.frombuffer(buf, hdt, 3 + nb, pos).astype(np.int64)
        p = pos + hs * (3 + nb)
        out = []
        for cs in head[3:]:
            out.append(zlib.decompress(buf[p:p + cs]))
            p += int(cs)
        return b"".join(out)
    text = b"".join(text.split())
    if not compressed:
        # header and data are one base64 stream
        first, _ = _b64(text, pos, hs)
        n = int(np.frombuffer(first, hdt, 1)[0])
        both, _ = _b64(text, pos, hs + n)
        return both[hs:hs + n]
    # compressed: the block table and the blocks are separate base64 streams
    first, _ = _b64(text, pos, 3 * hs)
    nb = int(np.frombuffer(first, hdt, 1)[0])
    hbytes, p = _b64(text, pos, hs * (3 + nb))
    head = np.frombuffer(hbytes, hdt, 3 + nb).astype(np.int64)
    data, _ = _b64(text, p, int(head[3:].sum()))
    out, q = [], 0
    for cs in head[3:]:
        out.append(zlib.decompress(data[q:q + cs]))
        q += int(cs)
    return b"".join(out)


def _array(da, appended, app_enc, hs, compressed, byte_order) -> np.ndarray:
    dt = np.dtype(_TYPES[da.get("type")]).newbyteorder("<" if byte_order == "LittleEndian" else ">")
    ncomp = int(da.get("NumberOfComponents", "1"))
    fmt = da.get("format", "ascii")
    if fmt == "ascii":
        a = np.array((da.text or "").split(), dtype=np.float64 if dt.kind == "f" else np.int64).astype(dt)
    elif fmt == "binary":
        a = np.frombuffer(_block(None, (da.text or "").encode(), 0, hs, compressed), dt)
    elif fmt == "appended":
        off = int(da.get("offset"))
        if app_enc == "raw":
            a = np.frombuffer(_block(appended, None, off, hs, compressed), dt)
        else:
            a = np.frombuffer(_block(None, appended, off, hs, compressed), dt)
    else:
        raise ValueError(f"DataArray format {fmt!r}")
    a = a.astype(dt.newbyteorder("="))
    return a.reshape(-1, ncomp) if ncomp > 1 else a


def read(path: str) -> VtkData:
    """Reads a VTK XML UnstructuredGrid (.vtu) or PolyData (.vtp) file (first piece)."""
    raw = open(path, "rb").read()
    appended, app_enc = None, None
    i = raw.find(b"<AppendedData")
    if i >= 0:
        j = raw.find(b">", i)
        tag = raw[i:j + 1].decode()
        app_enc = "raw" if 'encoding="raw"' in tag else "base64"
        us = raw.find(b"_", j)
        end = raw.rfind(b"</AppendedData>")
        appended = raw[us + 1:end]
        if app_enc == "base64":
            appended = appended.strip()
        raw = raw[:i] + b"</VTKFile>"
    root = ET.fromstring(raw)
    kind = root.get("type")
    if kind not in ("UnstructuredGrid", "PolyData"):
        raise ValueError(f"{path}: VTK file type {kind!r} is not a mesh")
    hs = 8 if root.get("header_type", "UInt32") == "UInt64" else 4
    compressed = root.get("compressor") is not None
    if compressed and root.get("compressor") != "vtkZLibDataCompressor":
        raise ValueError(f"{path}: compressor {root.get('compressor')} is not supported")
    bo = root.get("byte_order", "LittleEndian")
    piece = root.find(kind).find("Piece")

    def arr(da):
        return _array(da, appended, app_enc, hs, compressed, bo)

    pts = arr(piece.find("Points").find("DataArray")).astype(np.float64).reshape(-1, 3)
    if kind == "UnstructuredGrid":
        cells = piece.find("Cells")
    else:
        cells = piece.find("Polys")
        if cells is None or int(piece.get("NumberOfPolys", "0")) == 0:
            raise ValueError(f"{path}: a face file needs polygons")
    named = {da.get("Name"): da for da in cells.findall("DataArray")}
    conn = arr(named["connectivity"]).astype(np.int64)
    offs = arr(named["offsets"]).astype(np.int64)
    types = arr(named["types"]).astype(np.int64) if "types" in named else None
    pdn, cdn = piece.find("PointData"), piece.find("CellData")
    pd = {da.get("Name"): arr(da) for da in (pdn if pdn is not None else [])}
    cd = {da.get("Name"): arr(da) for da in (cdn if cdn is not None else [])}
    return VtkData(kind, pts, conn, offs, types, pd, cd)


# ------------------------------------------------------------------ writing --
def _payload(a: np.ndarray, hs: int, compress: bool) -> bytes:
    data = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes()
    hdt = "<u4" if hs == 4 else "<u8"
    if not compress:
        return np.array([len(data)], hdt).tobytes() + data
    bs = 1 << 15
    blocks = [data[k:k + bs] for k in range(0, len(data), bs)] or [b""]
    comp = [zlib.compress(b) for b in blocks]
    head = [len(blocks), bs, len(blocks[-1])] + [len(c) for c in comp]
    return np.array(head, hdt).tobytes() + b"".join(comp)


def write(path: str, kind: str, points: np.ndarray, cells: np.ndarray, cell_type: int, point_data=None,
          cell_data=None, fmt: str = "appended", compress: bool = True):
    """Writes one piece: fmt "appended" (raw appended data, UInt64 headers,
    zlib blocks when compress) or "ascii"."""
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    n_pts, n_cells, size = points.shape[0], cells.shape[0], cells.shape[1]
    arrays = []  # (section, name, array)
    arrays.append(("Points", None, np.ascontiguousarray(points, dtype=np.float64)))
    sec = "Cells" if kind == "UnstructuredGrid" else "Polys"
    arrays.append((sec, "connectivity", cells.ravel()))
    arrays.append((sec, "offsets", np.arange(1, n_cells + 1, dtype=np.int64) * size))
    if kind == "UnstructuredGrid":
        arrays.append((sec, "types", np.full(n_cells, cell_type, dtype=np.uint8)))
    for k, v in (point_data or {}).items():
        arrays.append(("PointData", k, np.asarray(v)))
    for k, v in (cell_data or {}).items():
        arrays.append(("CellData", k, np.asarray(v)))
    appended = bytearray()
    body = {}
    for sec_, name, a in arrays:
        ncomp = a.shape[1] if a.ndim > 1 else 1
        attrs = f'type="{_NAMES[a.dtype]}"' + (f' Name="{name}"' if name else "") + \
                f' NumberOfComponents="{ncomp}"'
        if fmt == "ascii":
            txt = " ".join(repr(float(x)) if a.dtype.kind == "f" else str(int(x)) for x in a.ravel())
            line = f'<DataArray {attrs} format="ascii">{txt}</DataArray>'
        else:
            line = f'<DataArray {attrs} format="appended" offset="{len(appended)}"/>'
            appended += _payload(a, 8, compress)
        body.setdefault(sec_, []).append(line)
    counts = (f'NumberOfPoints="{n_pts}" NumberOfCells="{n_cells}"' if kind == "UnstructuredGrid"
              else f'NumberOfPoints="{n_pts}" NumberOfPolys="{n_cells}"')
    head = '<?xml version="1.0"?>\n<VTKFile type="%s" version="1.0" byte_order="LittleEndian" header_type="UInt64"%s>\n' % (
        kind, ' compressor="vtkZLibDataCompressor"' if (compress and fmt != "ascii") else "")
    out = [head, f"<{kind}>\n<Piece {counts}>\n"]
    for sec_ in ("PointData", "CellData", "Points", sec):
        out.append(f"<{sec_}>\n" + "\n".join(body.get(sec_, [])) + f"\n</{sec_}>\n")
    out.append(f"</Piece>\n</{kind}>\n")
    with open(path, "wb") as f:
        f.write("".join(out).encode())
        if fmt != "ascii":
            f.write(b'<AppendedData encoding="raw">\n_')
            f.write(bytes(appended))
            f.write(b"\n</AppendedData>\n")
        f.write(b"</VTKFile>\n")


def write_vtu(path, points, tets, point_data=None, cell_data=None, **kw):
    write(path, "UnstructuredGrid", points, tets, VTK_TETRA, point_data, cell_data, **kw)


def write_vtp(path, points, tris, point_data=None, cell_data=None, **kw):
    write(path, "PolyData", points, tris, VTK_TRIANGLE, point_data, cell_data, **kw)


# ---------------------------------------------------------- mesh and faces --
@dataclass
class Face:
    """A face of the volume mesh (faceType): node ids gN, TRI3 connectivity in
    volume node ids, parent elements gE."""
    name: str
    gN: np.ndarray
    IEN: np.ndarray
    gE: np.ndarray


def read_face(path: str, name: str | None = None) -> Face:
    """A .vtp face with the load_vtp conventions (1-based GlobalNodeID/ElementID)."""
    d = read(path)
    if "GlobalNodeID" not in d.point_data or "GlobalElementID" not in d.cell_data:
        raise ValueError(f"{path}: a face needs GlobalNodeID point data and GlobalElementID cell data")
    gN = d.point_data["GlobalNodeID"].astype(np.int64) - 1
    tris = gN[d.cells(3)]
    gE = d.cell_data["GlobalElementID"].astype(np.int64) - 1
    return Face(name or os.path.splitext(os.path.basename(path))[0], gN.astype(np.int32), tris.astype(np.int32),
                gE.astype(np.int32))


def load_mesh(vtu_path: str, faces: dict, dirichlet=(), node_flags=None):
    """The fluid path's mesh from a SimVascular model: the TET4 volume mesh
    (.vtu) and its faces (.vtp, name -> path). Returns (Pipe, {name: Face});
    the Pipe's FSILS faces are the `dirichlet` ones in the given order (their
    node ids), its block-CSR pattern is built from the connectivity, and
    node_flags (face name -> flag bit) marks the nodes of named faces."""
    from .workload import Pipe, pattern
    d = read(vtu_path)
    if d.types is not None and np.any(d.types != VTK_TETRA):
        raise ValueError(f"{vtu_path}: only linear tetrahedra (VTK type 10) are on the fluid path")
    IEN = np.ascontiguousarray(d.cells(4), dtype=np.int32)
    x = np.ascontiguousarray(d.points)
    fcs = {k: read_face(v, k) for k, v in faces.items()}
    for f in fcs.values():
        if f.gN.size and (f.gN.min() < 0 or f.gN.max() >= x.shape[0]):
            raise ValueError(f"face {f.name}: GlobalNodeID outside the mesh")
        if f.gE.size and (f.gE.min() < 0 or f.gE.max() >= IEN.shape[0]):
            raise ValueError(f"face {f.name}: GlobalElementID outside the mesh")
    rowPtr, colPtr = pattern(x.shape[0], IEN)
    lo, hi = x.min(0), x.max(0)
    flags = np.zeros(x.shape[0], np.int32)
    for name, bit in (node_flags or {}).items():
        flags[fcs[name].gN] |= bit
    p = Pipe(0, 0, x, IEN, flags, rowPtr, colPtr, float(0.5 * max(hi[:2] - lo[:2])), float(hi[2] - lo[2]))
    p.faces = [(k, fcs[k].gN) for k in dirichlet]
    return p, fcs


def write_result(path: str, pipe, Yn, An=None, fmt: str = "appended", compress: bool = True):
    """A result file as the reference writes one (vtk_xml.cpp write_vtus):
    Velocity (3), Pressure and, when given, Acceleration point data, with
    GlobalNodeID / GlobalElementID (1-based)."""
    Y = np.asarray(Yn.cpu().numpy() if hasattr(Yn, "cpu") else Yn, dtype=np.float64).reshape(-1, 4)
    pdata = {"GlobalNodeID": np.arange(1, pipe.nNo + 1, dtype=np.int32),
             "Velocity": np.ascontiguousarray(Y[:, :3]), "Pressure": np.ascontiguousarray(Y[:, 3])}
    if An is not None:
        A = np.asarray(An.cpu().numpy() if hasattr(An, "cpu") else An, dtype=np.float64).reshape(-1, 4)
        pdata["Acceleration"] = np.ascontiguousarray(A[:, :3])
    cdata = {"GlobalElementID": np.arange(1, pipe.nEl + 1, dtype=np.int32)}
    write_vtu(path, pipe.x, pipe.IEN, pdata, cdata, fmt=fmt, compress=compress)
