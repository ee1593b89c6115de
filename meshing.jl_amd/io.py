"""Data formats on either side of the extraction: NRRD volumes in, PLY
meshes out.

The reference's examples load a CT volume from an NRRD file
(examples/nrrd.jl:8-11,21; docs/src/examples.md:14-27) and save meshes as
PLY (etc/output_comparison.jl:41-42) through FileIO/NRRD.jl/MeshIO, which
are not part of Meshing.jl.  These readers and writers give the same
hand-off here:

* ``read_nrrd(path)`` -> a dense array indexed ``[x, y, z]`` with x fastest
  in memory (NRRD lists the fastest axis first, Julia's column-major
  order), ready for ``isosurface``; ``NrrdInfo`` carries the header fields
  (spacing, origin) for building the X, Y, Z axes.
* ``write_ply(path, vertices, faces)`` / ``read_ply(path)``: triangle meshes
  as returned by ``isosurface`` (1-based faces in memory, 0-based in the
  file, as the PLY format has them).

Host code (numpy, the standard library's gzip/bz2): file formats are byte
streams, and the volume goes to the device once, in ``isosurface``.
"""
import bz2
import gzip
import os
from dataclasses import dataclass, field as dc_field

import numpy as np

__all__ = ["NrrdInfo", "read_nrrd", "write_nrrd", "read_ply", "write_ply"]

# NRRD type names (http://teem.sourceforge.net/nrrd/format.html#type) -> numpy
_NRRD_TYPES = {}
for _names, _dt in (
        (("signed char", "int8", "int8_t"), np.int8),
        (("uchar", "unsigned char", "uint8", "uint8_t"), np.uint8),
        (("short", "short int", "signed short", "signed short int", "int16", "int16_t"),
         np.int16),
        (("ushort", "unsigned short", "unsigned short int", "uint16", "uint16_t"), np.uint16),
        (("int", "signed int", "int32", "int32_t"), np.int32),
        (("uint", "unsigned int", "uint32", "uint32_t"), np.uint32),
        (("longlong", "long long", "long long int", "signed long long",
          "signed long long int", "int64", "int64_t"), np.int64),
        (("ulonglong", "unsigned long long", "unsigned long long int", "uint64", "uint64_t"),
         np.uint64),
        (("float",), np.float32),
        (("double",), np.float64)):
    for _n in _names:
        _NRRD_TYPES[_n] = _dt
_NUMPY_TO_NRRD = {np.dtype(np.int8): "int8", np.dtype(np.uint8): "uint8",
                  np.dtype(np.int16): "int16", np.dtype(np.uint16): "uint16",
                  np.dtype(np.int32): "int32", np.dtype(np.uint32): "uint32",
                  np.dtype(np.int64): "int64", np.dtype(np.uint64): "uint64",
                  np.dtype(np.float32): "float", np.dtype(np.float64): "double"}


@dataclass
class NrrdInfo:
    """Header of an NRRD file: ``sizes`` fastest axis first, the stored
    element type, and the optional geometry fields (per-axis spacing, or
    space directions and origin) as given."""

    sizes: tuple
    dtype: np.dtype
    encoding: str
    spacings: tuple = None
    space_directions: list = None
    space_origin: tuple = None
    fields: dict = dc_field(default_factory=dict)

    def axes(self):
        """X, Y, Z ranges ``(first, last)`` from the origin and the axis
        spacing (space directions, else spacings; unit spacing from 0
        otherwise), for ``isosurface(A, method, X, Y, Z)``."""
        out = []
        for k, n in enumerate(self.sizes[:3]):
            o = self.space_origin[k] if self.space_origin else 0.0
            if self.space_directions and self.space_directions[k] is not None:
                d = float(np.linalg.norm(self.space_directions[k]))
            elif self.spacings and np.isfinite(self.spacings[k]):
                d = float(self.spacings[k])
            else:
                d = 1.0
            out.append((float(o), float(o) + d * (n - 1)))
        return tuple(out)


def _vec(s):
    s = s.strip()
    if s == "none":
        return None
    return tuple(float(t) for t in s.strip("()").split(","))


def _parse_header(f):
    magic = f.readline()
    if not magic.startswith(b"NRRD000"):
        raise ValueError("not an NRRD file (missing NRRD000x magic)")
    fields, keyvals = {}, {}
    while True:
        line = f.readline()
        if not line:
            break  # header without data (detached)
        line = line.decode("latin-1").rstrip("\r\n")
        if line == "":
            break
        if line.startswith("#"):
            continue
        if ":=" in line:
            k, v = line.split(":=", 1)
            keyvals[k] = v
            continue
        if ": " not in line:
            raise ValueError(f"bad NRRD header line {line!r}")
        k, v = line.split(": ", 1)
        fields[k.strip().lower()] = v.strip()
    return fields, keyvals


def read_nrrd(path, dtype=None):
    """Read an NRRD volume; returns ``(array, info)``.

    The array is indexed ``[x, y, z]`` (fastest file axis first) with x
    fastest in memory.  ``dtype`` converts it (e.g. ``numpy.float32``: the
    GPU path takes Float32 and Float64 fields; integer volumes such as CT
    data convert exactly up to 2^24).  Encodings raw, gzip, bzip2 and
    ascii/text; little and big endian; detached data files; line and byte
    skips."""
    with open(path, "rb") as f:
        fields, keyvals = _parse_header(f)
        attached_at = f.tell()
    if "type" not in fields or "sizes" not in fields or "encoding" not in fields:
        raise ValueError("NRRD header needs type, sizes and encoding")
    t = fields["type"].lower()
    if t not in _NRRD_TYPES:
        raise ValueError(f"unsupported NRRD type {fields['type']!r}")
    sizes = tuple(int(s) for s in fields["sizes"].split())
    dim = int(fields.get("dimension", len(sizes)))
    if dim != len(sizes):
        raise ValueError("NRRD dimension does not match sizes")
    enc = fields["encoding"].lower()
    enc = {"txt": "ascii", "text": "ascii", "gz": "gzip", "bz2": "bzip2"}.get(enc, enc)
    dt = np.dtype(_NRRD_TYPES[t])
    if dt.itemsize > 1:
        endian = fields.get("endian", "little").lower()
        dt = dt.newbyteorder("<" if endian == "little" else ">")
    count = int(np.prod(sizes)) if sizes else 0
    datafile = fields.get("data file", fields.get("datafile"))
    if datafile is not None:
        if datafile.startswith("LIST") or "%" in datafile.split()[0]:
            raise ValueError("multi-file NRRD data is not supported")
        dpath = datafile if os.path.isabs(datafile) else os.path.join(
            os.path.dirname(os.path.abspath(path)), datafile)
        with open(dpath, "rb") as g:
            blob = g.read()
    else:
        with open(path, "rb") as g:
            g.seek(attached_at)
            blob = g.read()
    # line skip, then byte skip (byte skip applies to the decoded stream
    # for compressed encodings; -1 means "the data end the stream", raw only)
    for _ in range(int(fields.get("line skip", fields.get("lineskip", 0)))):
        nl = blob.find(b"\n")
        blob = blob[nl + 1:] if nl >= 0 else b""
    bskip = int(fields.get("byte skip", fields.get("byteskip", 0)))
    if enc == "gzip":
        blob = gzip.decompress(blob)
    elif enc == "bzip2":
        blob = bz2.decompress(blob)
    elif enc not in ("raw", "ascii"):
        raise ValueError(f"unsupported NRRD encoding {fields['encoding']!r}")
    if enc == "ascii":
        a = np.array(blob.decode("latin-1").split()[:count], dtype=dt.newbyteorder("="))
    else:
        nbytes = count * dt.itemsize
        if bskip == -1:
            if enc != "raw":
                raise ValueError("byte skip -1 is only defined for raw data")
            blob = blob[len(blob) - nbytes:]
        else:
            blob = blob[bskip:]
        if len(blob) < nbytes:
            raise ValueError(f"NRRD data too short: {len(blob)} of {nbytes} bytes")
        a = np.frombuffer(blob[:nbytes], dtype=dt).astype(dt.newbyteorder("="))
    if a.size != count:
        raise ValueError(f"NRRD data has {a.size} of {count} values")
    a = a.reshape(sizes, order="F")
    if dtype is not None:
        a = np.asfortranarray(a.astype(dtype))
    dirs = None
    if "space directions" in fields:
        dirs = [_vec(v) for v in fields["space directions"].replace(") (", ")  (").split()]
    spac = None
    if "spacings" in fields:
        spac = tuple(float("nan") if v == "nan" else float(v) for v in fields["spacings"].split())
    origin = _vec(fields["space origin"]) if "space origin" in fields else None
    info = NrrdInfo(sizes, np.dtype(_NRRD_TYPES[t]), enc, spac, dirs, origin,
                    dict(fields, **{"key/values": keyvals}))
    return a, info


def write_nrrd(path, array, encoding="raw", endian="little", spacings=None):
    """Write a 3-D (or any-D) array indexed with the fastest axis first
    (``[x, y, z]``) as an attached-data NRRD file."""
    a = np.asarray(array)
    if a.dtype not in _NUMPY_TO_NRRD:
        raise ValueError(f"no NRRD type for {a.dtype}")
    if encoding not in ("raw", "gzip", "ascii"):
        raise ValueError("encoding must be raw, gzip or ascii")
    lines = ["NRRD0004", f"type: {_NUMPY_TO_NRRD[a.dtype]}", f"dimension: {a.ndim}",
             "sizes: " + " ".join(str(s) for s in a.shape), f"encoding: {encoding}"]
    if a.dtype.itemsize > 1 and encoding != "ascii":
        lines.append(f"endian: {endian}")
    if spacings is not None:
        lines.append("spacings: " + " ".join(repr(float(s)) for s in spacings))
    flat = a.reshape(-1, order="F")
    if encoding == "ascii":
        body = (" ".join(repr(v.item()) for v in flat) + "\n").encode()
    else:
        body = flat.astype(a.dtype.newbyteorder("<" if endian == "little" else ">")).tobytes()
        if encoding == "gzip":
            body = gzip.compress(body)
    with open(path, "wb") as f:
        f.write(("\n".join(lines) + "\n\n").encode())
        f.write(body)


def _host(x):
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ImportError:
        pass
    return np.asarray(x)


def write_ply(path, vertices, faces, binary=True):
    """Write a triangle mesh as PLY: ``vertices`` (n, 3) Float32 or Float64
    (PLY ``float`` / ``double``), ``faces`` (m, 3) 1-based indices as
    ``isosurface`` returns them (written 0-based, as ``int`` lists, or
    ``uint`` beyond 2^31 vertices).  CUDA tensors are copied to the host."""
    V = _host(vertices)
    F = _host(faces).astype(np.int64, copy=False)
    if V.ndim != 2 or V.shape[1] != 3 or F.ndim != 2 or F.shape[1] != 3:
        raise ValueError("vertices and faces must be (n, 3) arrays")
    if V.dtype not in (np.float32, np.float64):
        raise ValueError("vertex coordinates must be Float32 or Float64")
    nv = V.shape[0]
    if F.size and (F.min() < 1 or F.max() > nv):
        raise ValueError("face index outside 1..nverts")
    if nv >= 2 ** 32:
        raise ValueError("PLY indices are 32-bit")
    it, idt = ("int", np.int32) if nv < 2 ** 31 else ("uint", np.uint32)
    vt = "float" if V.dtype == np.float32 else "double"
    fmt = "binary_little_endian" if binary else "ascii"
    header = (f"ply\nformat {fmt} 1.0\ncomment written by meshing.jl_amd\n"
              f"element vertex {nv}\nproperty {vt} x\nproperty {vt} y\nproperty {vt} z\n"
              f"element face {F.shape[0]}\nproperty list uchar {it} vertex_indices\n"
              "end_header\n")
    with open(path, "wb") as f:
        f.write(header.encode())
        if binary:
            f.write(np.ascontiguousarray(V, dtype=V.dtype.newbyteorder("<")).tobytes())
            rec = np.zeros(F.shape[0], dtype=[("n", "u1"),
                                              ("i", np.dtype(idt).newbyteorder("<"), (3,))])
            rec["n"] = 3
            rec["i"] = F - 1
            f.write(rec.tobytes())
        else:
            for p in V:
                f.write((" ".join(repr(float(c)) if V.dtype == np.float64 else
                                  repr(float(np.float32(c))) for c in p) + "\n").encode())
            for q in F - 1:
                f.write(f"3 {q[0]} {q[1]} {q[2]}\n".encode())


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2",
              "int16": "i2", "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4",
              "uint": "u4", "uint32": "u4", "float": "f4", "float32": "f4", "double": "f8",
              "float64": "f8"}


def read_ply(path):
    """Read a triangle-mesh PLY (ascii or binary, either endianness) whose
    vertex element starts with x, y, z and whose face element is one list
    of triangles; returns ``(vertices, faces)`` with 1-based int64 faces,
    the vertex type as stored."""
    with open(path, "rb") as f:
        if f.readline().strip() != b"ply":
            raise ValueError("not a PLY file")
        fmt, elems = None, []
        while True:
            line = f.readline()
            if not line:
                raise ValueError("PLY header without end_header")
            tok = line.decode("latin-1").split()
            if not tok or tok[0] in ("comment", "obj_info"):
                continue
            if tok[0] == "end_header":
                break
            if tok[0] == "format":
                fmt = tok[1]
            elif tok[0] == "element":
                elems.append((tok[1], int(tok[2]), []))
            elif tok[0] == "property":
                elems[-1][2].append(tuple(tok[1:]))
        body = f.read()
    if fmt not in ("ascii", "binary_little_endian", "binary_big_endian"):
        raise ValueError(f"unknown PLY format {fmt!r}")
    end = ">" if fmt == "binary_big_endian" else "<"
    V = F = None
    pos = 0
    words = body.decode("latin-1").split() if fmt == "ascii" else None
    for name, n, props in elems:
        lists = [p for p in props if p[0] == "list"]
        if not lists:
            names = [p[-1] for p in props]
            if fmt == "ascii":
                k = len(props)
                vals = np.array(words[pos:pos + n * k], dtype=np.float64).reshape(n, k)
                pos += n * k
                rec = {nm: vals[:, j].astype(_PLY_TYPES[props[j][0]]) for j, nm in enumerate(names)}
            else:
                dt = np.dtype([(nm, end + _PLY_TYPES[p[0]]) for nm, p in zip(names, props)])
                arr = np.frombuffer(body, dtype=dt, count=n, offset=pos)
                pos += n * dt.itemsize
                rec = {nm: arr[nm] for nm in names}
            if name == "vertex":
                if names[:3] != ["x", "y", "z"]:
                    raise ValueError("vertex element must start with x, y, z")
                V = np.stack([rec["x"], rec["y"], rec["z"]], axis=1).astype(
                    rec["x"].dtype.newbyteorder("="))
        else:
            if len(props) != 1:
                raise ValueError("face element must hold one list")
            _, ct, it, _ = lists[0]
            if fmt == "ascii":
                rows = []
                for _ in range(n):
                    c = int(words[pos])
                    if c != 3:
                        raise ValueError("only triangle faces are supported")
                    rows.append(words[pos + 1:pos + 4])
                    pos += 1 + c
                idx = np.array(rows, dtype=np.int64).reshape(n, 3)
            else:
                dt = np.dtype([("n", end + _PLY_TYPES[ct]), ("i", end + _PLY_TYPES[it], (3,))])
                arr = np.frombuffer(body, dtype=dt, count=n, offset=pos)
                if n and np.any(arr["n"] != 3):
                    raise ValueError("only triangle faces are supported")
                pos += n * dt.itemsize
                idx = arr["i"].astype(np.int64)
            if name == "face":
                F = idx + 1
    if V is None or F is None:
        raise ValueError("PLY file needs vertex and face elements")
    return V, F
