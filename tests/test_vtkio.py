"""The .pvti / .vti reader (particle_tracking/example_kitchensink.py:7-36):
files written here in every supported encoding read back to the cell arrays
they were made from, for scalar and 3-component data and for a parallel
image split into pieces."""
import base64
import struct
import zlib

import numpy as np
import pytest

from turbulence_tracing_amd.vtkio import pvti_readin


def cells_flat(a):
    """(ncells, ncomp) in VTK order (x fastest) from a[ix, iy, iz(, c)]."""
    a = a if a.ndim == 4 else a[..., None]
    return a.reshape(-1, a.shape[3], order="F")


def vti_text(a, ext, spacing, fmt, compress=False, header="UInt32", order="<"):
    ncomp = 1 if a.ndim == 3 else a.shape[3]
    data = cells_flat(a).astype(order + "f8").tobytes()
    h = "I" if header == "UInt32" else "Q"
    if compress:
        comp = zlib.compress(data)
        head = struct.pack(order + h * 4, 1, len(data), len(data), len(comp))
        payload_raw = head + comp
    else:
        payload_raw = struct.pack(order + h, len(data)) + data
    attrs = f'type="ImageData" version="1.0" byte_order="{"LittleEndian" if order == "<" else "BigEndian"}" ' \
            f'header_type="{header}"' + (' compressor="vtkZLibDataCompressor"' if compress else "")
    extent = " ".join(str(v) for v in ext)
    if fmt == "ascii":
        body = " ".join(repr(float(v)) for v in cells_flat(a).ravel())
        da = f'<DataArray type="Float64" Name="rnec" NumberOfComponents="{ncomp}" format="ascii">{body}</DataArray>'
        tail = b""
    elif fmt == "binary":
        if compress:
            txt = base64.b64encode(head).decode() + base64.b64encode(comp).decode()
        else:
            txt = base64.b64encode(payload_raw).decode()
        da = f'<DataArray type="Float64" Name="rnec" NumberOfComponents="{ncomp}" format="binary">{txt}</DataArray>'
        tail = b""
    else:
        da = f'<DataArray type="Float64" Name="rnec" NumberOfComponents="{ncomp}" format="appended" offset="0"/>'
        tail = b'<AppendedData encoding="raw">_' + payload_raw + b"</AppendedData>"
    head_xml = (f'<?xml version="1.0"?><VTKFile {attrs}><ImageData WholeExtent="{extent}" Origin="0 0 0" '
                f'Spacing="{" ".join(str(s) for s in spacing)}"><Piece Extent="{extent}"><PointData/>'
                f'<CellData Scalars="rnec">{da}</CellData></Piece></ImageData>')
    return head_xml.encode() + tail + b"</VTKFile>"


@pytest.mark.parametrize("fmt,compress,header,order", [
    ("ascii", False, "UInt32", "<"), ("binary", False, "UInt32", "<"), ("binary", True, "UInt64", "<"),
    ("appended", False, "UInt64", "<"), ("appended", True, "UInt32", "<"), ("appended", False, "UInt32", ">")])
@pytest.mark.parametrize("ncomp", [1, 3])
def test_vti_round_trip(tmp_path, fmt, compress, header, order, ncomp):
    rs = np.random.RandomState(ncomp)
    a = rs.randn(5, 4, 3) if ncomp == 1 else rs.randn(5, 4, 3, 3)
    p = tmp_path / "cube.vti"
    p.write_bytes(vti_text(a, (0, 5, 0, 4, 0, 3), (0.5, 0.25, 0.125), fmt, compress, header, order))
    img, dim, spacing = pvti_readin(str(p))
    np.testing.assert_array_equal(img, a)
    assert dim == a.shape
    np.testing.assert_array_equal(spacing, [0.5, 0.25, 0.125])


def test_pvti_assembles_pieces(tmp_path):
    rs = np.random.RandomState(7)
    a = rs.randn(6, 4, 5)
    srcs = []
    for i, (x0, x1) in enumerate([(0, 2), (2, 6)]):
        name = f"piece_{i}.vti"
        (tmp_path / name).write_bytes(vti_text(a[x0:x1], (x0, x1, 0, 4, 0, 5), (1e-5, 1e-5, 2e-5), "appended",
                                               compress=(i == 1)))
        srcs.append((name, (x0, x1, 0, 4, 0, 5)))
    pieces = "".join(f'<Piece Extent="{" ".join(map(str, e))}" Source="{n}"/>' for n, e in srcs)
    (tmp_path / "cube.pvti").write_text(
        '<?xml version="1.0"?><VTKFile type="PImageData" version="1.0" byte_order="LittleEndian">'
        '<PImageData WholeExtent="0 6 0 4 0 5" GhostLevel="0" Origin="0 0 0" Spacing="1e-05 1e-05 2e-05">'
        '<PCellData Scalars="rnec"><PDataArray type="Float64" Name="rnec"/></PCellData>' + pieces +
        "</PImageData></VTKFile>")
    img, dim, spacing = pvti_readin(str(tmp_path / "cube.pvti"))
    np.testing.assert_array_equal(img, a)
    np.testing.assert_array_equal(spacing, [1e-5, 1e-5, 2e-5])
