"""Reader for VTK XML image data (.vti) and parallel image data (.pvti).

``pvti_readin(filename)`` follows the reference's reader
(particle_tracking/example_kitchensink.py:7-36), which uses VTK's
vtkXMLPImageDataReader: the first cell-data array of the image, reshaped
Fortran-order to (nx-1, ny-1, nz-1[, ncomp]) cells, with the point
dimensions and spacing.  VTK is not needed: the XML is parsed with
ElementTree and the arrays decoded with numpy.  Supported encodings: ascii,
inline base64 ("binary"), and appended data (raw or base64), each
uncompressed or zlib-compressed (vtkZLibDataCompressor), little or big
endian, UInt32 or UInt64 headers.  A .pvti assembles its pieces by extent.
"""
import base64
import os
import re
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_TYPES = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
          "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}


def _split_appended(raw):
    """(xml text with the appended block emptied, appended bytes after '_')."""
    m = re.search(rb"<AppendedData[^>]*>", raw)
    if m is None:
        return raw.decode("utf-8", "replace"), b""
    start = raw.index(b"_", m.end()) + 1
    end = raw.rindex(b"</AppendedData>")
    xml = raw[:m.end()] + raw[end:]
    return xml.decode("utf-8", "replace"), raw[start:end]


def _read_blocks(buf, pos, hdr, compressed, b64):
    """Decode one array starting at buf[pos:] (bytes or base64 text)."""
    hsize = np.dtype(hdr).itemsize
    if b64:
        text = buf[pos:] if isinstance(buf, (bytes, bytearray)) else buf[pos:].encode()
        text = re.sub(rb"\s", b"", text)
        if not compressed:
            head = base64.b64decode(text[:4 * ((hsize + 2) // 3)])
            n = int(np.frombuffer(head[:hsize], dtype=hdr)[0])
            return base64.b64decode(text[:4 * ((hsize + n + 2) // 3)])[hsize:hsize + n]
        head = base64.b64decode(text[:4 * ((3 * hsize + 2) // 3)])
        nblocks = int(np.frombuffer(head[:hsize], dtype=hdr)[0])
        hlen = (3 + nblocks) * hsize
        hb64 = 4 * ((hlen + 2) // 3)
        header = np.frombuffer(base64.b64decode(text[:hb64])[:hlen], dtype=hdr)
        data = base64.b64decode(text[hb64:hb64 + 4 * ((int(header[3:].sum()) + 2) // 3)])
        return _inflate(header, data)
    if not compressed:
        n = int(np.frombuffer(buf[pos:pos + hsize], dtype=hdr)[0])
        return buf[pos + hsize:pos + hsize + n]
    nblocks = int(np.frombuffer(buf[pos:pos + hsize], dtype=hdr)[0])
    header = np.frombuffer(buf[pos:pos + (3 + nblocks) * hsize], dtype=hdr)
    start = pos + (3 + nblocks) * hsize
    return _inflate(header, buf[start:start + int(header[3:].sum())])


def _inflate(header, data):
    out, off = [], 0
    for c in header[3:]:
        out.append(zlib.decompress(data[off:off + int(c)]))
        off += int(c)
    return b"".join(out)


def _data_array(arr, root, appended):
    dt = np.dtype(_TYPES[arr.attrib["type"]])
    if root.attrib.get("byte_order", "LittleEndian") == "BigEndian":
        dt = dt.newbyteorder(">")
    hdr = np.dtype({"UInt32": "u4", "UInt64": "u8"}[root.attrib.get("header_type", "UInt32")])
    if dt.byteorder == ">":
        hdr = hdr.newbyteorder(">")
    compressed = "compressor" in root.attrib
    fmt = arr.attrib.get("format", "ascii")
    if fmt == "ascii":
        vals = np.array(arr.text.split(), dtype=dt)
    elif fmt == "binary":
        vals = np.frombuffer(_read_blocks(arr.text.strip(), 0, hdr, compressed, True), dtype=dt)
    else:  # appended
        enc = root.find("AppendedData").attrib.get("encoding", "raw")
        vals = np.frombuffer(_read_blocks(appended, int(arr.attrib["offset"]), hdr, compressed, enc == "base64"),
                             dtype=dt)
    return vals.reshape(-1, int(arr.attrib.get("NumberOfComponents", "1")))


def _read_vti(path):
    """(cells as (ncells, ncomp), cell extent ranges, spacing)."""
    with open(path, "rb") as f:
        xml, appended = _split_appended(f.read())
    root = ET.fromstring(xml)
    img = root.find("ImageData")
    piece = img.find("Piece")
    ext = [int(v) for v in piece.attrib.get("Extent", img.attrib["WholeExtent"]).split()]
    spacing = [float(v) for v in img.attrib.get("Spacing", "1 1 1").split()]
    cd = piece.find("CellData")
    arr = cd.findall("DataArray")[0]
    return _data_array(arr, root, appended), ext, spacing


def pvti_readin(filename):
    """The first cell-data array of a .pvti (or .vti) image: returns (img,
    dim, spacing) like particle_tracking/example_kitchensink.py:7-36, with
    img[ix, iy, iz(, c)] over cells, dim = img.shape and spacing the image's
    (dx, dy, dz)."""
    with open(filename, "rb") as f:
        raw = f.read()
    xml, _ = _split_appended(raw)
    root = ET.fromstring(xml)
    if root.attrib.get("type") == "ImageData":
        cells, ext, spacing = _read_vti(filename)
        whole = ext
        pieces = [(cells, ext)]
    else:
        pim = root.find("PImageData")
        whole = [int(v) for v in pim.attrib["WholeExtent"].split()]
        spacing = [float(v) for v in pim.attrib.get("Spacing", "1 1 1").split()]
        pieces = []
        base = os.path.dirname(os.path.abspath(filename))
        for p in pim.findall("Piece"):
            cells, ext, _ = _read_vti(os.path.join(base, p.attrib["Source"]))
            pieces.append((cells, [int(v) for v in p.attrib["Extent"].split()]))
    dim = [whole[1] - whole[0] + 1, whole[3] - whole[2] + 1, whole[5] - whole[4] + 1]
    ncomp = pieces[0][0].shape[1]
    out = np.zeros([d - 1 for d in dim] + [ncomp], dtype=pieces[0][0].dtype)
    for cells, ext in pieces:
        shp = [ext[1] - ext[0], ext[3] - ext[2], ext[5] - ext[4]]
        block = cells.reshape(shp + [ncomp], order="F")
        out[ext[0] - whole[0]:ext[1] - whole[0], ext[2] - whole[2]:ext[3] - whole[2],
            ext[4] - whole[4]:ext[5] - whole[4]] = block
    img = out if ncomp > 1 else out[..., 0]
    return img, img.shape, np.array(spacing)
