"""CPU checks of the data formats either side of the path: NRRD volumes in
(examples/nrrd.jl:8-11,21) and PLY meshes out (etc/output_comparison.jl:41-42).
Round trips through every encoding and endianness, a hand-written NRRD
header with known values, detached data and skips, and the error paths."""
import gzip

import numpy as np
import pytest


@pytest.fixture(scope="module")
def io(pkg):
    from importlib import import_module
    return import_module(pkg.__name__ + ".io")


@pytest.mark.parametrize("dtype", [np.int16, np.uint8, np.float32, np.float64, np.int32])
@pytest.mark.parametrize("encoding", ["raw", "gzip", "ascii"])
@pytest.mark.parametrize("endian", ["little", "big"])
def test_nrrd_round_trip(io, tmp_path, dtype, encoding, endian):
    rng = np.random.default_rng(1)
    a = (rng.standard_normal((7, 5, 3)) * 100).astype(dtype)
    p = tmp_path / "v.nrrd"
    io.write_nrrd(p, a, encoding=encoding, endian=endian, spacings=(0.5, 1.0, 2.0))
    b, info = io.read_nrrd(p)
    assert b.dtype == a.dtype and b.shape == a.shape and np.array_equal(a, b)
    assert b.flags.f_contiguous
    assert info.sizes == (7, 5, 3) and info.spacings == (0.5, 1.0, 2.0)
    assert info.axes() == ((0.0, 3.0), (0.0, 4.0), (0.0, 4.0))
    c, _ = io.read_nrrd(p, dtype=np.float32)
    assert c.dtype == np.float32 and np.array_equal(c, a.astype(np.float32))


def test_nrrd_hand_written(io, tmp_path):
    # x fastest: values 0..23 laid out as sizes 4 3 2
    vals = np.arange(24, dtype=">i2")
    hdr = ("NRRD0005\n# a comment\ntype: short\ndimension: 3\nsizes: 4 3 2\n"
           "endian: big\nencoding: raw\nspace: left-posterior-superior\n"
           "space directions: (0.5,0,0) (0,0.5,0) (0,0,2)\n"
           "space origin: (-1,-2,-3)\nmodality:=CT\n\n")
    p = tmp_path / "h.nrrd"
    p.write_bytes(hdr.encode() + vals.tobytes())
    a, info = io.read_nrrd(p)
    assert a.shape == (4, 3, 2) and a[1, 0, 0] == 1 and a[0, 1, 0] == 4 and a[0, 0, 1] == 12
    assert a[3, 2, 1] == 23
    assert info.space_origin == (-1.0, -2.0, -3.0)
    assert info.axes() == ((-1.0, 0.5), (-2.0, -1.0), (-3.0, -1.0))
    assert info.fields["key/values"] == {"modality": "CT"}


def test_nrrd_detached_and_skips(io, tmp_path):
    vals = np.arange(8, dtype="<f4")
    (tmp_path / "d.raw").write_bytes(b"junk line\n" + b"XXXX" + vals.tobytes())
    (tmp_path / "d.nhdr").write_text("NRRD0004\ntype: float\ndimension: 3\nsizes: 2 2 2\n"
                                     "encoding: raw\nendian: little\nline skip: 1\n"
                                     "byte skip: 4\ndata file: d.raw\n")
    a, _ = io.read_nrrd(tmp_path / "d.nhdr")
    assert np.array_equal(a.reshape(-1, order="F"), vals)
    (tmp_path / "e.raw").write_bytes(b"HEADERSTUFF" + vals.tobytes())
    (tmp_path / "e.nhdr").write_text("NRRD0004\ntype: float\nsizes: 2 2 2\nencoding: raw\n"
                                     "byte skip: -1\ndata file: e.raw\n")
    a, _ = io.read_nrrd(tmp_path / "e.nhdr")
    assert np.array_equal(a.reshape(-1, order="F"), vals)
    # gzip: byte skip applies after decompression
    (tmp_path / "g.raw.gz").write_bytes(gzip.compress(b"ab" + vals.tobytes()))
    (tmp_path / "g.nhdr").write_text("NRRD0004\ntype: float\nsizes: 2 2 2\nencoding: gzip\n"
                                     "byte skip: 2\ndata file: g.raw.gz\n")
    a, _ = io.read_nrrd(tmp_path / "g.nhdr")
    assert np.array_equal(a.reshape(-1, order="F"), vals)


def test_nrrd_errors(io, tmp_path):
    p = tmp_path / "bad.nrrd"
    p.write_bytes(b"P6\n")
    with pytest.raises(ValueError):
        io.read_nrrd(p)
    p.write_bytes(b"NRRD0004\ntype: float\nsizes: 4 4 4\nencoding: raw\n\n" + b"\0" * 10)
    with pytest.raises(ValueError):
        io.read_nrrd(p)
    p.write_bytes(b"NRRD0004\ntype: quaternion\nsizes: 1\nencoding: raw\n\n")
    with pytest.raises(ValueError):
        io.read_nrrd(p)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("binary", [True, False])
def test_ply_round_trip(io, oracle, tmp_path, dtype, binary):
    from helpers import sphere_field
    # iso of the field's type, so the mesh keeps it (Julia's promotion)
    V, F = oracle.isosurface(sphere_field(20, 0.6, dtype), "mc", iso=dtype(0))
    assert V.dtype == dtype
    p = tmp_path / "m.ply"
    io.write_ply(p, V, F, binary=binary)
    V2, F2 = io.read_ply(p)
    assert V2.dtype == V.dtype and np.array_equal(V2.view(np.uint8), V.view(np.uint8))
    assert F2.dtype == np.int64 and np.array_equal(F2, F)
    head = p.read_bytes().split(b"end_header")[0].decode()
    assert ("double x" if dtype == np.float64 else "float x") in head
    assert "property list uchar int vertex_indices" in head


def test_ply_big_endian_and_errors(io, tmp_path):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    body = V.astype(">f4").tobytes() + np.array([3], "u1").tobytes() + \
        np.array([0, 1, 2], ">i4").tobytes()
    p = tmp_path / "b.ply"
    p.write_bytes(b"ply\nformat binary_big_endian 1.0\nelement vertex 3\nproperty float x\n"
                  b"property float y\nproperty float z\nelement face 1\n"
                  b"property list uchar int vertex_indices\nend_header\n" + body)
    V2, F2 = io.read_ply(p)
    assert np.array_equal(V2, V) and F2.tolist() == [[1, 2, 3]]
    with pytest.raises(ValueError):
        io.write_ply(tmp_path / "x.ply", V, np.array([[1, 2, 4]]))
    with pytest.raises(ValueError):
        io.write_ply(tmp_path / "x.ply", V.astype(np.float16), np.array([[1, 2, 3]]))


def test_nrrd_to_mesh_to_ply(io, oracle, tmp_path):
    # the reference example's flow on the CPU oracle: an Int16 CT-like
    # volume read as Float32 (exact), extracted, saved
    from helpers import sphere_field
    ct = np.round(sphere_field(24, 0.5, np.float64) * 1000).astype(np.int16)
    io.write_nrrd(tmp_path / "ct.nrrd", ct, encoding="gzip")
    a, info = io.read_nrrd(tmp_path / "ct.nrrd", dtype=np.float32)
    V, F = oracle.isosurface(a, "mt", iso=np.float32(100))
    io.write_ply(tmp_path / "ct.ply", V, F)
    V2, F2 = io.read_ply(tmp_path / "ct.ply")
    assert np.array_equal(V2, V) and np.array_equal(F2, F) and len(F) > 0
