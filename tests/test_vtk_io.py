"""VTK XML mesh and result files (svfsiplus_amd/vtk_io.py): the encodings VTK
writes, hand-encoded here from the file-format rules (inline base64 with the
header and data in one stream; zlib-compressed blocks with a separately
encoded block table; appended raw and base64 data; UInt32 and UInt64
headers), the 1-based face ids of load_vtp, and a SimVascular-style model
directory round trip. No file written by VTK itself is available (the
reference's meshes are LFS pointers and VTK is not installed), so agreement
with VTK's own writer rests on these format rules."""
import base64
import os
import zlib

import numpy as np
import pytest

from svfsiplus_amd import _lib, vtk_io as V, workload as W

# one tetrahedron and one extra point
PTS = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1]], dtype=np.float64)
TET = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], dtype=np.int64)
VEL = np.arange(15, dtype=np.float64).reshape(5, 3) / 7.0


def _hdr(values, hs):
    return np.array(values, dtype="<u4" if hs == 4 else "<u8").tobytes()


def _inline(a, hs, compress):
    data = a.astype(a.dtype.newbyteorder("<")).tobytes()
    if not compress:
        return base64.b64encode(_hdr([len(data)], hs) + data).decode()
    bs = 40  # several blocks, a partial last one
    blocks = [data[k:k + bs] for k in range(0, len(data), bs)]
    comp = [zlib.compress(b) for b in blocks]
    head = _hdr([len(blocks), bs, len(blocks[-1])] + [len(c) for c in comp], hs)
    return base64.b64encode(head).decode() + base64.b64encode(b"".join(comp)).decode()


def _vtu_inline(hs, compress):
    ht = "UInt32" if hs == 4 else "UInt64"
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    arr = lambda a, t, name, nc: (f'<DataArray type="{t}" Name="{name}" NumberOfComponents="{nc}" format="binary">'
                                  f'{_inline(a, hs, compress)}</DataArray>')
    return f'''<?xml version="1.0"?>
<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="{ht}"{comp}>
 <UnstructuredGrid><Piece NumberOfPoints="5" NumberOfCells="2">
  <PointData>{arr(VEL, "Float64", "Velocity", 3)}</PointData>
  <CellData>{arr(np.array([7, 9], np.int32), "Int32", "ModelRegionID", 1)}</CellData>
  <Points>{arr(PTS, "Float64", "Points", 3)}</Points>
  <Cells>{arr(TET.ravel(), "Int64", "connectivity", 1)}{arr(np.array([4, 8], np.int64), "Int64", "offsets", 1)}
   {arr(np.array([10, 10], np.uint8), "UInt8", "types", 1)}</Cells>
 </Piece></UnstructuredGrid>
</VTKFile>
'''


@pytest.mark.parametrize("hs", [4, 8])
@pytest.mark.parametrize("compress", [False, True])
def test_inline_base64(tmp_path, hs, compress):
    f = tmp_path / "m.vtu"
    f.write_text(_vtu_inline(hs, compress))
    d = V.read(str(f))
    assert np.array_equal(d.points, PTS) and np.array_equal(d.cells(4), TET)
    assert np.array_equal(d.point_data["Velocity"], VEL)
    assert d.cell_data["ModelRegionID"].tolist() == [7, 9] and d.types.tolist() == [10, 10]


@pytest.mark.parametrize("enc", ["raw", "base64"])
@pytest.mark.parametrize("compress", [False, True])
def test_appended(tmp_path, enc, compress):
    hs = 4
    arrays = [("PointData", "Velocity", "Float64", 3, VEL), ("Points", "Points", "Float64", 3, PTS),
              ("Cells", "connectivity", "Int64", 1, TET.ravel()), ("Cells", "offsets", "Int64", 1,
                                                                  np.array([4, 8], np.int64)),
              ("Cells", "types", "UInt8", 1, np.array([10, 10], np.uint8))]
    blob, xml = b"", {}
    for sec, name, t, nc, a in arrays:
        data = a.astype(a.dtype.newbyteorder("<")).tobytes()
        if compress:
            comp = zlib.compress(data)
            head = _hdr([1, len(data), len(data), len(comp)], hs)
            piece = (head + comp) if enc == "raw" else (base64.b64encode(head) + base64.b64encode(comp))
        else:
            piece = (_hdr([len(data)], hs) + data) if enc == "raw" else base64.b64encode(_hdr([len(data)], hs) + data)
        xml.setdefault(sec, []).append(f'<DataArray type="{t}" Name="{name}" NumberOfComponents="{nc}" '
                                       f'format="appended" offset="{len(blob)}"/>')
        blob += piece
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    head = (f'<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
            f'header_type="UInt32"{comp}>\n<UnstructuredGrid><Piece NumberOfPoints="5" NumberOfCells="2">'
            + "".join(f"<{s}>{''.join(v)}</{s}>" for s, v in xml.items()) +
            f'</Piece></UnstructuredGrid>\n<AppendedData encoding="{enc}">\n   _')
    f = tmp_path / "a.vtu"
    f.write_bytes(head.encode() + blob + b"\n  </AppendedData>\n</VTKFile>\n")
    d = V.read(str(f))
    assert np.array_equal(d.points, PTS) and np.array_equal(d.cells(4), TET)
    assert np.array_equal(d.point_data["Velocity"], VEL)


@pytest.mark.parametrize("fmt,compress", [("appended", True), ("appended", False), ("ascii", False)])
def test_model_directory_round_trip(tmp_path, fmt, compress):
    """A SimVascular-style model written from the seeded pipe (volume .vtu,
    face .vtp files with 1-based GlobalNodeID/GlobalElementID) loads back
    into the same mesh, pattern and faces."""
    p = W.make_pipe(4)
    fi, ge, inner = W.outlet_faces(p)
    surf = tmp_path / "mesh-surfaces"
    surf.mkdir()
    V.write_vtu(str(tmp_path / "mesh-complete.mesh.vtu"), p.x, p.IEN,
                {"GlobalNodeID": np.arange(1, p.nNo + 1, dtype=np.int32)},
                {"GlobalElementID": np.arange(1, p.nEl + 1, dtype=np.int32)}, fmt=fmt, compress=compress)
    nodes = np.unique(fi)
    local = np.searchsorted(nodes, fi)
    V.write_vtp(str(surf / "outlet.vtp"), p.x[nodes], local, {"GlobalNodeID": (nodes + 1).astype(np.int32)},
                {"GlobalElementID": (ge + 1).astype(np.int32)}, fmt=fmt, compress=compress)
    q, faces = V.load_mesh(str(tmp_path / "mesh-complete.mesh.vtu"), {"outlet": str(surf / "outlet.vtp")},
                           node_flags={"outlet": _lib.NODE_OUTLET})
    assert np.array_equal(q.x, p.x) and np.array_equal(q.IEN, p.IEN)
    assert np.array_equal(q.rowPtr, p.rowPtr) and np.array_equal(q.colPtr, p.colPtr)
    f = faces["outlet"]
    assert np.array_equal(f.gN, nodes) and np.array_equal(f.IEN, fi) and np.array_equal(f.gE, ge)
    # the flags rebuild the same outlet TRI3 faces
    fi2, ge2, inner2 = W.outlet_faces(q)
    assert np.array_equal(fi2, fi) and np.array_equal(ge2, ge) and np.array_equal(inner2, inner)
    # results
    Y = np.random.default_rng(3).standard_normal((p.nNo, 4))
    V.write_result(str(tmp_path / "result_001.vtu"), p, Y, Y * 2, fmt=fmt, compress=compress)
    r = V.read(str(tmp_path / "result_001.vtu"))
    assert np.array_equal(r.point_data["Velocity"], Y[:, :3]) and np.array_equal(r.point_data["Pressure"], Y[:, 3])
    assert np.array_equal(r.point_data["Acceleration"], 2 * Y[:, :3])
    assert r.point_data["GlobalNodeID"][0] == 1 and r.cell_data["GlobalElementID"][-1] == p.nEl


def test_rejects_other_cells_and_bad_ids(tmp_path):
    V.write(str(tmp_path / "t.vtu"), "UnstructuredGrid", PTS, np.array([[0, 1, 2]]), V.VTK_TRIANGLE)
    with pytest.raises(ValueError):
        V.load_mesh(str(tmp_path / "t.vtu"), {})
    V.write_vtu(str(tmp_path / "m.vtu"), PTS, TET)
    V.write_vtp(str(tmp_path / "f.vtp"), PTS[:3], np.array([[0, 1, 2]]), {"GlobalNodeID": np.array([1, 2, 99])},
                {"GlobalElementID": np.array([1])})
    with pytest.raises(ValueError):
        V.load_mesh(str(tmp_path / "m.vtu"), {"f": str(tmp_path / "f.vtp")})
