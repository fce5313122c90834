"""The .vti loader (pyvr_amd/dataloaders/vtk_loader.py), against the
reference's own test expectations (tests/test_dataloaders/test_vtk_loader.py)
on its example files (copied as fixtures to tests/golden/example_data), the
value range those files declare in their headers, and round trips of every
storage variant through a small writer here."""

import base64
import re
import struct
import zlib
from pathlib import Path

import numpy as np
import pytest

from pyvr_amd.dataloaders import load_vtk_volume
from pyvr_amd.dataloaders.vtk_loader import _array_values, read_vti
from pyvr_amd.datasets import compute_normal_volume

DATA = Path(__file__).parent / "golden" / "example_data"


# -- the reference's expectations ------------------------------------------

@pytest.mark.parametrize("name,shape", [("hydrogen", (128, 128, 128)), ("fuel", (64, 64, 64))])
def test_example_files(name, shape):
    v = load_vtk_volume(DATA / f"{name}.vti")
    assert v.name == f"{name}.vti(Scalars_)"
    assert v.shape == shape and v.data.dtype == np.float32
    assert v.has_normals and v.normals.shape == shape + (3,)
    assert 0.0 <= v.data.min() < 0.1 and 0.9 < v.data.max() <= 1.0
    assert np.allclose(v.center, 0.0, atol=1e-6)
    assert np.isclose(np.max(np.abs(v.max_bounds)), 1.0, atol=1e-6)
    assert np.allclose((v.max_bounds - v.min_bounds) / (v.max_bounds - v.min_bounds)[0], 1.0)
    assert np.array_equal(v.normals, compute_normal_volume(v.data))


@pytest.mark.parametrize("name", ["hydrogen", "fuel"])
def test_raw_values_match_declared_range(name):
    """The decoded array spans exactly the RangeMin/RangeMax the file states."""
    text = (DATA / f"{name}.vti").read_bytes()[:2000].decode("latin-1")
    lo = float(re.search(r'RangeMin="([^"]+)"', text).group(1))
    hi = float(re.search(r'RangeMax="([^"]+)"', text).group(1))
    dims, _, arrays, fmt, appended = read_vti(DATA / f"{name}.vti")
    vals = _array_values(arrays["Scalars_"][0], fmt, appended)
    assert vals.size == np.prod(dims)
    assert float(vals.min()) == lo and float(vals.max()) == hi


def test_errors_and_paths():
    with pytest.raises(FileNotFoundError, match="not found"):
        load_vtk_volume("nonexistent.vti")
    with pytest.raises(ValueError) as e:
        load_vtk_volume(DATA / "hydrogen.vti", scalars_name="InvalidName")
    assert "not found" in str(e.value) and "Available arrays" in str(e.value) and "Scalars_" in str(e.value)
    assert load_vtk_volume(str(DATA / "fuel.vti")).shape == (64, 64, 64)


# -- round trips -------------------------------------------------------------

def write_vti(path, data_xyz, spacing=(1, 1, 1), mode="appended-b64", zlib_blocks=0, header="UInt32",
              big=False, name="Scalars_", ncomp=1, extra=None, split_header=False):
    """Minimal VTK ImageData writer: data_xyz is indexed [x, y, z]
    (components last when ncomp > 1).  As VTK does, an uncompressed array's
    header and data are one base64 stream, a compressed array's header is
    encoded on its own; ``split_header`` encodes it on its own always."""
    nx, ny, nz = data_xyz.shape[:3]
    e = ">" if big else "<"
    hfmt = "Q" if header == "UInt64" else "I"
    flat = np.ascontiguousarray(np.moveaxis(data_xyz, (0, 1, 2), (2, 1, 0)))   # z slowest ... x fastest
    vtype = {np.dtype("f4"): "Float32", np.dtype("f8"): "Float64", np.dtype("u1"): "UInt8",
             np.dtype("i2"): "Int16", np.dtype("u2"): "UInt16", np.dtype("i4"): "Int32"}[flat.dtype]
    raw = flat.astype(flat.dtype.newbyteorder(e)).tobytes()

    def blocks():
        if zlib_blocks:
            bs = max(1, -(-len(raw) // zlib_blocks))
            parts = [raw[i:i + bs] for i in range(0, len(raw), bs)]
            comp = [zlib.compress(p) for p in parts]
            head = struct.pack(f"{e}{3 + len(comp)}{hfmt}", len(comp), bs, len(parts[-1]), *[len(c) for c in comp])
            return head, b"".join(comp)
        return struct.pack(f"{e}{hfmt}", len(raw)), raw

    def b64(head, payload):
        if zlib_blocks or split_header:
            return base64.b64encode(head) + base64.b64encode(payload)
        return base64.b64encode(head + payload)

    comp_attr = ' compressor="vtkZLibDataCompressor"' if zlib_blocks else ""
    order = "BigEndian" if big else "LittleEndian"
    nc = f' NumberOfComponents="{ncomp}"' if ncomp != 1 else ""
    arrays, appended = [], b""
    specs = [(name, vtype)] + ([(extra, "Float32")] if extra else [])
    for nm, ty in specs:
        if mode == "ascii":
            body = " ".join(repr(float(v)) if flat.dtype.kind == "f" else str(int(v)) for v in flat.ravel())
            arrays.append(f'<DataArray type="{ty}" Name="{nm}"{nc} format="ascii">{body}</DataArray>')
        elif mode == "binary":
            head, payload = blocks()
            txt = b64(head, payload).decode()
            arrays.append(f'<DataArray type="{ty}" Name="{nm}"{nc} format="binary">{txt}</DataArray>')
        else:
            head, payload = blocks()
            chunk = b64(head, payload) if mode == "appended-b64" else head + payload
            arrays.append(f'<DataArray type="{ty}" Name="{nm}"{nc} format="appended" offset="{len(appended)}"/>')
            appended += chunk
    xml = (f'<?xml version="1.0"?>\n<VTKFile type="ImageData" version="0.1" byte_order="{order}" '
           f'header_type="{header}"{comp_attr}>\n<ImageData WholeExtent="0 {nx - 1} 0 {ny - 1} 0 {nz - 1}" '
           f'Origin="0 0 0" Spacing="{spacing[0]} {spacing[1]} {spacing[2]}">\n<Piece Extent="0 {nx - 1} 0 {ny - 1} '
           f'0 {nz - 1}">\n<PointData Scalars="{name}">\n' + "\n".join(arrays) + "\n</PointData>\n</Piece>\n</ImageData>\n")
    out = xml.encode()
    if mode.startswith("appended"):
        enc = "base64" if mode == "appended-b64" else "raw"
        out += f'<AppendedData encoding="{enc}">\n_'.encode() + appended + b"\n</AppendedData>\n"
    out += b"</VTKFile>\n"
    Path(path).write_bytes(out)


@pytest.mark.parametrize("mode", ["ascii", "binary", "appended-b64", "appended-raw"])
@pytest.mark.parametrize("zb", [0, 1, 3])
@pytest.mark.parametrize("header,big", [("UInt32", False), ("UInt64", True)])
def test_round_trip(tmp_path, mode, zb, header, big):
    if mode == "ascii" and (zb or big or header == "UInt64"):
        pytest.skip("ascii arrays have no blocks or byte order")
    rng = np.random.default_rng(zb + 10 * big)
    d = rng.random((5, 7, 3)).astype(np.float32) * 100 - 20
    p = tmp_path / "v.vti"
    write_vti(p, d, mode=mode, zlib_blocks=zb, header=header, big=big, extra="Other")
    dims, _, arrays, fmt, appended = read_vti(p)
    assert dims == (5, 7, 3) and set(arrays) == {"Scalars_", "Other"}
    vals = _array_values(arrays["Scalars_"][0], fmt, appended)
    assert np.array_equal(vals.reshape(3, 7, 5), np.moveaxis(d, (0, 1, 2), (2, 1, 0)))
    v = load_vtk_volume(p)
    assert v.shape == (3, 7, 5)                       # (nz, ny, nx)
    ref = np.moveaxis(d, (0, 1, 2), (2, 1, 0))
    assert np.array_equal(v.data, (ref - ref.min()) / (ref.max() - ref.min()))


@pytest.mark.parametrize("mode", ["binary", "appended-b64"])
@pytest.mark.parametrize("header", ["UInt32", "UInt64"])
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_uncompressed_header_encoded_apart(tmp_path, mode, header, n):
    """Both base64 layouts of an uncompressed array, at payload sizes of
    every remainder modulo 3 (4-byte values: 4, 8, 12, 20 bytes)."""
    d = np.arange(n, dtype=np.float32).reshape(n, 1, 1) - 1.5
    ref = np.moveaxis(d, (0, 1, 2), (2, 1, 0)).ravel()
    for split in (False, True):
        p = tmp_path / f"s{int(split)}.vti"
        write_vti(p, d, mode=mode, header=header, split_header=split)
        dims, _, arrays, fmt, appended = read_vti(p)
        assert np.array_equal(_array_values(arrays["Scalars_"][0], fmt, appended), ref)


@pytest.mark.parametrize("dtype", ["u1", "i2", "u2", "i4", "f8"])
def test_integer_and_double_arrays(tmp_path, dtype):
    d = (np.arange(4 * 3 * 2).reshape(4, 3, 2) * 3).astype(dtype)
    p = tmp_path / "i.vti"
    write_vti(p, d, mode="appended-b64", zlib_blocks=2)
    v = load_vtk_volume(p)
    ref = np.moveaxis(d, (0, 1, 2), (2, 1, 0)).astype(np.float32)
    assert np.array_equal(v.data, (ref - ref.min()) / (ref.max() - ref.min()))


def test_aspect_constant_and_multicomponent(tmp_path):
    p = tmp_path / "a.vti"
    write_vti(p, np.ones((4, 6, 2), np.float32), spacing=(2.0, 0.5, 3.0))
    v = load_vtk_volume(p)
    phys = np.array([4 * 2.0, 6 * 0.5, 2 * 3.0], np.float32)
    assert np.allclose(v.max_bounds, phys / phys.max()) and np.allclose(v.min_bounds, -v.max_bounds)
    assert not v.data.any()                           # constant field -> zeros
    write_vti(p, np.zeros((3, 3, 3, 3), np.float32), ncomp=3)
    with pytest.raises(ValueError, match="Multi-component"):
        load_vtk_volume(p)
    write_vti(p, np.ones((2, 2, 2), np.float32), spacing=(1, 0, 1))
    with pytest.raises(ValueError, match="spacing"):
        load_vtk_volume(p)


@pytest.mark.parametrize("zb", [0, 2])
def test_raw_payload_ending_in_whitespace_bytes(tmp_path, zb):
    """Raw appended data whose last bytes are 0x0a / 0x20 / 0x09 (values
    chosen so) loads intact: the payload is read by header sizes only."""
    d = np.zeros((2, 2, 2), np.uint32)
    d.flat[-1] = 0x0A200A20
    d.flat[-2] = 0x090D0A20
    p = tmp_path / "w.vti"
    write_vti(p, d.astype(np.int32), mode="appended-raw", zlib_blocks=zb)
    dims, _, arrays, fmt, appended = read_vti(p)
    vals = _array_values(arrays["Scalars_"][0], fmt, appended)
    assert np.array_equal(vals, np.moveaxis(d.astype(np.int32), (0, 1, 2), (2, 1, 0)).ravel())
