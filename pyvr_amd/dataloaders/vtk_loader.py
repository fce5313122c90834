"""``load_vtk_volume``: VTK XML ImageData (``.vti``) files as ``Volume``
objects, with the behaviour of the reference loader
(``pyvr/dataloaders/vtk_loader.py:15-146``): scalars reshaped to
(nz, ny, nx), normalised to [0, 1] in float32 (all zeros for a range below
1e-9), bounds centred at the origin with the longest physical side mapped to
[-1, 1] (physical size = dimensions x spacing), name ``"<file>(<array>)"``,
normals computed.

The reference reads the file through VTK, which is not a dependency here;
this module parses the XML format with the standard library: ``ascii``,
inline ``binary`` and ``appended`` (raw or base64) data arrays, with or
without ``vtkZLibDataCompressor`` blocks, UInt32 or UInt64 headers, either
byte order.
"""

import base64
import re
import struct
import xml.etree.ElementTree as ET
import zlib
from pathlib import Path
from typing import Union

import numpy as np

from ..volume import Volume

_DTYPES = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
           "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}


class _Format:
    def __init__(self, root):
        self.endian = "<" if root.get("byte_order", "LittleEndian") == "LittleEndian" else ">"
        self.header = "u8" if root.get("header_type", "UInt32") == "UInt64" else "u4"
        comp = root.get("compressor")
        if comp not in (None, "", "vtkZLibDataCompressor"):
            raise ValueError(f"Unsupported VTK compressor: {comp}")
        self.zlib = comp == "vtkZLibDataCompressor"

    @property
    def hsize(self) -> int:
        return 8 if self.header == "u8" else 4

    def ints(self, raw: bytes, n: int):
        return struct.unpack(f"{self.endian}{n}{'Q' if self.header == 'u8' else 'I'}", raw[: n * self.hsize])


def _block_payload(fmt: _Format, read_b64, raw: bytes = None):
    """Payload bytes of one data array.  ``read_b64(char_from, nchars)``
    decodes base64 text (a compressed array's header and blocks are encoded
    separately); ``raw`` is the unencoded appended stream from the array's
    offset."""
    hs = fmt.hsize
    if fmt.zlib:
        if raw is not None:
            nb = fmt.ints(raw, 1)[0]
            head = fmt.ints(raw, 3 + nb)
            pos, blocks = hs * (3 + nb), []
            for size in head[3:]:
                blocks.append(zlib.decompress(raw[pos:pos + size]))
                pos += size
            return b"".join(blocks)
        nb = fmt.ints(read_b64(0, 4 * ((hs + 2) // 3)), 1)[0]
        hlen = hs * (3 + nb)
        head_chars = 4 * ((hlen + 2) // 3)
        head = fmt.ints(read_b64(0, head_chars), 3 + nb)
        comp = read_b64(head_chars, None)
        pos, blocks = 0, []
        for size in head[3:]:
            blocks.append(zlib.decompress(comp[pos:pos + size]))
            pos += size
        return b"".join(blocks)
    if raw is not None:
        n = fmt.ints(raw, 1)[0]
        return raw[hs:hs + n]
    # VTK encodes an uncompressed array's header and data as one base64
    # stream; a header encoded on its own ends in padding within its quantum
    head_chars = 4 * ((hs + 2) // 3)
    head = read_b64(0, head_chars)
    if len(head) == hs:
        n = fmt.ints(head, 1)[0]
        return read_b64(head_chars, None)[:n]
    whole = read_b64(0, None)
    n = fmt.ints(whole, 1)[0]
    return whole[hs:hs + n]


def _b64_reader(text: str):
    text = re.sub(r"\s+", "", text)

    def read(start, nchars):
        chunk = text[start:] if nchars is None else text[start:start + nchars]
        return base64.b64decode(chunk)
    return read


def _array_values(array, fmt: _Format, appended) -> np.ndarray:
    dtype = np.dtype(fmt.endian + _DTYPES[array.get("type")])
    kind = array.get("format", "ascii")
    if kind == "ascii":
        return np.array((array.text or "").split(), dtype=dtype.newbyteorder("="))
    if kind == "binary":
        payload = _block_payload(fmt, _b64_reader(array.text or ""))
    elif kind == "appended":
        if appended is None:
            raise ValueError("DataArray refers to missing AppendedData")
        encoding, data = appended
        off = int(array.get("offset", "0"))
        if encoding == "base64":
            text = re.sub(r"\s+", "", data.decode("ascii", errors="strict"))
            payload = _block_payload(fmt, _b64_reader(text[off:]))
        else:
            payload = _block_payload(fmt, None, data[off:])
    else:
        raise ValueError(f"Unsupported DataArray format: {kind}")
    return np.frombuffer(payload, dtype=dtype)


def _split_appended(content: bytes):
    """(xml bytes with the appended payload removed, (encoding, payload))."""
    m = re.search(rb"<AppendedData\b([^>]*)>", content)
    if m is None:
        return content, None
    enc = re.search(rb'encoding="([^"]*)"', m.group(1))
    encoding = enc.group(1).decode() if enc else "raw"
    start = content.index(b"_", m.end()) + 1
    end = content.rindex(b"</AppendedData>")
    # raw payloads are read by the sizes their headers give, so bytes after
    # the last array (whitespace before the closing tag) are never used; no
    # stripping, which could cut data bytes that happen to be whitespace
    payload = content[start:end]
    xml = content[:m.end()] + content[end:]
    return xml, (encoding, payload)


def read_vti(file_path: Union[str, Path]):
    """(dims (nx, ny, nz), spacing (sx, sy, sz), {array name: (values,
    components)}) of a VTK ImageData file's point data."""
    content = Path(file_path).read_bytes()
    xml, appended = _split_appended(content)
    root = ET.fromstring(xml)
    if root.tag != "VTKFile" or root.get("type") != "ImageData":
        raise ValueError("Not a VTK ImageData file")
    fmt = _Format(root)
    image = root.find("ImageData")
    ext = [int(v) for v in image.get("WholeExtent").split()]
    dims = (ext[1] - ext[0] + 1, ext[3] - ext[2] + 1, ext[5] - ext[4] + 1)
    spacing = tuple(float(v) for v in image.get("Spacing", "1 1 1").split())
    arrays = {}
    piece = image.find("Piece")
    pdata = piece.find("PointData") if piece is not None else None
    if pdata is not None:
        for a in pdata.findall("DataArray"):
            arrays[a.get("Name")] = (a, int(a.get("NumberOfComponents", "1")))
    return dims, spacing, arrays, fmt, appended


def load_vtk_volume(file_path: Union[str, Path], scalars_name: str = "Scalars_") -> Volume:
    """Load a VTK ImageData (.vti) file as a ``Volume`` (see the module doc).

    Raises ``FileNotFoundError`` for a missing file and ``ValueError`` for
    invalid dimensions or spacing, a missing array (listing the available
    ones) or a multi-component array."""
    file_path = Path(file_path)
    if not file_path.exists():
        raise FileNotFoundError(f"VTK file not found: {file_path}")
    dims, spacing, arrays, fmt, appended = read_vti(file_path)
    if any(d <= 0 for d in dims):
        raise ValueError(f"Invalid dimensions: {dims}")
    if any(s <= 0 for s in spacing):
        raise ValueError(f"Invalid spacing (must be positive): {spacing}")
    if scalars_name not in arrays:
        raise ValueError(f"Scalar array '{scalars_name}' not found in {file_path}. "
                         f"Available arrays: {list(arrays)}")
    array, ncomp = arrays[scalars_name]
    if ncomp != 1:
        raise ValueError(f"Multi-component scalars not supported. "
                         f"Array '{scalars_name}' has {ncomp} components, expected 1.")
    values = _array_values(array, fmt, appended)
    n = dims[0] * dims[1] * dims[2]
    if values.size != n:
        raise ValueError(f"Array '{scalars_name}' holds {values.size} values, expected {n}")
    data = values.reshape((dims[2], dims[1], dims[0])).astype(np.float32)
    lo, hi = data.min(), data.max()
    if hi - lo < 1e-9:
        normalized = np.zeros_like(data)
    else:
        normalized = (data - lo) / (hi - lo)
    physical = np.array([dims[0] * spacing[0], dims[1] * spacing[1], dims[2] * spacing[2]], dtype=np.float32)
    scale = 2.0 / np.max(physical)
    half = physical * scale / 2.0
    volume = Volume(data=normalized, normals=None, min_bounds=-half, max_bounds=+half,
                    name=f"{file_path.name}({scalars_name})")
    volume.compute_normals()
    return volume
