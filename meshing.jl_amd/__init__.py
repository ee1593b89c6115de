"""meshing.jl_amd -- MI355X (gfx950) isosurface extraction with the interface of
Meshing.jl v0.7.0.

Host-side mirror of the reference's public API (src/Meshing.jl:9-11)::

    isosurface(sdf, method=MarchingCubes(), X=(-1, 1), Y=(-1, 1), Z=(-1, 1))
        -> (vertices, faces)

* ``sdf`` is indexed ``sdf[x, y, z]``, like a Julia ``Array{T,3}``: x is the
  fastest axis in memory.  A numpy array (any order; it is made Fortran
  ordered) is a host field; a CUDA ``torch.Tensor`` whose strides are
  ``(1, nx, nx*ny)`` (see :func:`device_field`) is a device field and gives
  device results.  Element types: float32 and float64.
* ``X, Y, Z`` are the axis arguments; only ``first`` and ``last`` matter
  (src/marching_cubes.jl:36-38).  Give a ``(first, last)`` pair, a Python
  ``range`` or a 1-D array.  Python ``int`` is Julia ``Int``, ``float`` is
  ``Float64`` and ``numpy.float32`` is ``Float32``; the same holds for
  ``iso`` and ``eps``, and these types decide the output precision exactly
  as Julia's promotion does (default ``iso=0.0`` gives Float64 vertices).
* ``vertices`` is ``(V, 3)``, ``faces`` is ``(F, 3)`` int64 and 1-based,
  both in the reference's order.

All computation runs in libmeshing_hip.so; nothing here computes a mesh.
"""
import ctypes
from dataclasses import dataclass
from numbers import Real

import numpy as np

from . import _capi
from ._capi import (MSH_DEVICE, MSH_FLOAT32, MSH_FLOAT64, MSH_HOST, MSH_INT64, MshField,
                    MshParams, context)

__all__ = ["isosurface", "isosurface_batch", "AbstractMeshingAlgorithm", "MarchingCubes",
           "MarchingTetrahedra", "NaiveSurfaceNets", "device_field", "generate_field", "count",
           "autotune", "autotune_batch", "set_kernel_family", "KERNEL_FAMILIES", "Server", "FieldSpec", "weld",
           "read_nrrd", "write_nrrd", "NrrdInfo", "read_ply", "write_ply"]

# emit kernel families, best first (see msh_autotune in include/meshing_hip.h)
KERNEL_FAMILIES = {3: "lanes along x, half-depth bricks", 2: "lanes along x",
                   1: "lanes along z", 0: "generic", -1: "default"}


class AbstractMeshingAlgorithm:
    """Abstract algorithm type (src/algorithmtypes.jl:10)."""

    method = -1


@dataclass(frozen=True)
class MarchingCubes(AbstractMeshingAlgorithm):
    """MarchingCubes(iso=0.0) (src/algorithmtypes.jl:23-25).  Vertices are not
    shared between cubes."""

    iso: Real = 0.0
    method = _capi.MSH_MARCHING_CUBES


@dataclass(frozen=True)
class MarchingTetrahedra(AbstractMeshingAlgorithm):
    """MarchingTetrahedra(iso=0.0, eps=1e-3) (src/algorithmtypes.jl:37-40).
    Vertices are shared; ``eps`` keeps them off voxel corners."""

    iso: Real = 0.0
    eps: Real = 1e-3
    method = _capi.MSH_MARCHING_TETRAHEDRA


@dataclass(frozen=True)
class NaiveSurfaceNets(AbstractMeshingAlgorithm):
    """NaiveSurfaceNets(iso=0.0): one vertex per cell crossing the surface,
    quads across crossing grid edges.  Named by the reference's README but
    without code in v0.7.0; the definition is this library's own."""

    iso: Real = 0.0
    method = _capi.MSH_NAIVE_SURFACE_NETS


def _kind(v, what):
    if isinstance(v, (bool, np.bool_)):
        raise TypeError(f"{what}: Bool is not supported")
    if isinstance(v, (int, np.integer)):
        return MSH_INT64
    if isinstance(v, np.float32):
        return MSH_FLOAT32
    if isinstance(v, (float, np.float64)):
        return MSH_FLOAT64
    raise TypeError(f"{what}: unsupported type {type(v).__name__} (Int, Float32, Float64)")


def _ends(R, what):
    if isinstance(R, range):
        if len(R) == 0:
            raise ValueError(f"{what}: empty range")
        return R[0], R[-1]
    if isinstance(R, np.ndarray):
        return R[0], R[-1]
    try:
        a, b = R
    except (TypeError, ValueError):
        raise TypeError(f"{what}: expected (first, last), a range or a 1-D array") from None
    return a, b


def _params(method, X, Y, Z):
    p = MshParams()
    p.method = method.method
    p.iso_kind = _kind(method.iso, "iso")
    if p.iso_kind == MSH_INT64:
        p.iso_int = int(method.iso)
    p.iso = float(method.iso)
    eps = getattr(method, "eps", 0)
    p.eps_kind = _kind(eps, "eps")
    if p.eps_kind == MSH_INT64:
        p.eps_int = int(eps)
    p.eps = float(eps)
    for k, (R, nm) in enumerate(((X, "X"), (Y, "Y"), (Z, "Z"))):
        a, b = _ends(R, nm)
        ka = _kind(a, nm)
        p.axis_kind[k] = ka
        p.first[k] = float(a)
        p.last[k] = float(b)
    return p


def _torch():
    import torch
    return torch


def _is_device_tensor(x):
    try:
        torch = _torch()
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def device_field(t):
    """Return a view of the CUDA tensor ``t`` (indexed [x, y, z]) whose memory
    is x-fastest, copying only if needed."""
    if t.dim() != 3:
        raise ValueError("sdf must be 3-dimensional")
    nx, ny, nz = t.shape
    if t.stride() == (1, nx, nx * ny):
        return t
    return t.permute(2, 1, 0).contiguous().permute(2, 1, 0)


def _field(sdf):
    if _is_device_tensor(sdf):
        torch = _torch()
        t = device_field(sdf)
        if t.dtype == torch.float32:
            dt = MSH_FLOAT32
        elif t.dtype == torch.float64:
            dt = MSH_FLOAT64
        else:
            raise TypeError(f"unsupported element type {t.dtype}; the GPU path takes "
                            "Float32 or Float64 fields")
        f = MshField(dt, MSH_DEVICE, *t.shape, t.data_ptr())
        return f, t, True
    a = np.asarray(sdf)
    if a.ndim != 3:
        raise ValueError("sdf must be 3-dimensional")
    if a.dtype == np.float32:
        dt = MSH_FLOAT32
    elif a.dtype == np.float64:
        dt = MSH_FLOAT64
    else:
        raise TypeError(f"unsupported element type {a.dtype}; the GPU path takes Float32 or "
                        "Float64 fields")
    a = np.asfortranarray(a)
    f = MshField(dt, MSH_HOST, *a.shape, a.ctypes.data)
    return f, a, False


def _split_args(method, args, X, Y, Z):
    # isosurface(A, args...) inserts MarchingCubes() when no algorithm is
    # given (src/isosurface.jl:30-32)
    if method is not None and not isinstance(method, AbstractMeshingAlgorithm):
        args = (method,) + tuple(args)
        method = None
    if method is None:
        method = MarchingCubes()
    axes = list(args) + [None] * (3 - len(args))
    if len(args) > 3:
        raise TypeError("too many positional arguments")
    X = axes[0] if axes[0] is not None else X
    Y = axes[1] if axes[1] is not None else Y
    Z = axes[2] if axes[2] is not None else Z
    return method, X, Y, Z


def count(sdf, method=None, *args, X=(-1, 1), Y=(-1, 1), Z=(-1, 1), device=0):
    """(number of vertices, number of faces) isosurface() would return."""
    method, X, Y, Z = _split_args(method, args, X, Y, Z)
    f, keep, on_dev = _field(sdf)
    p = _params(method, X, Y, Z)
    ctx = context(keep.device.index if on_dev else device)
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    with ctx.lock:
        if on_dev:
            ctx.check(_capi.lib().msh_set_stream(
                ctx.handle, ctypes.c_void_p(_torch().cuda.current_stream(keep.device).cuda_stream)))
        ctx.check(_capi.lib().msh_count(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                                        ctypes.byref(nv), ctypes.byref(nf)))
    return nv.value, nf.value


def isosurface(sdf, method=None, *args, X=(-1, 1), Y=(-1, 1), Z=(-1, 1), device=0):
    """Extract an isosurface.  Mirrors ``isosurface(A[, method][, X, Y, Z])``
    of Meshing.jl (src/isosurface.jl:30-32, src/marching_cubes.jl:27,
    src/marching_tetrahedra.jl:129)."""
    method, X, Y, Z = _split_args(method, args, X, Y, Z)
    if not isinstance(method, AbstractMeshingAlgorithm) or method.method < 0:
        raise TypeError(f"no isosurface method for {type(method).__name__}")
    f, keep, on_dev = _field(sdf)
    p = _params(method, X, Y, Z)
    L = _capi.lib()
    ok = ctypes.c_int32()
    ctx = context(keep.device.index if on_dev else device)
    nv, nf = ctypes.c_int64(), ctypes.c_int64()
    with ctx.lock:
        if on_dev:
            torch = _torch()
            stream = torch.cuda.current_stream(keep.device)
            ctx.check(L.msh_set_stream(ctx.handle, ctypes.c_void_p(stream.cuda_stream)))
        else:
            ctx.check(L.msh_set_stream(ctx.handle, None))
        ctx.check(L.msh_count(ctx.handle, ctypes.byref(f), ctypes.byref(p), ctypes.byref(nv),
                              ctypes.byref(nf)))
        ctx.check(L.msh_output_kind(ctypes.byref(f), ctypes.byref(p), ctypes.byref(ok)))
        if on_dev:
            tdt = torch.float32 if ok.value == MSH_FLOAT32 else torch.float64
            V = torch.empty((nv.value, 3), dtype=tdt, device=keep.device)
            F = torch.empty((nf.value, 3), dtype=torch.int64, device=keep.device)
            ctx.check(L.msh_extract(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                                    ctypes.c_void_p(V.data_ptr()), ctypes.c_void_p(F.data_ptr()),
                                    MSH_DEVICE))
        else:
            odt = np.float32 if ok.value == MSH_FLOAT32 else np.float64
            V = np.empty((nv.value, 3), dtype=odt)
            F = np.empty((nf.value, 3), dtype=np.int64)
            ctx.check(L.msh_extract(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                                    ctypes.c_void_p(V.ctypes.data), ctypes.c_void_p(F.ctypes.data),
                                    MSH_HOST))
    return V, F


def isosurface_batch(fields, methods=None, axes=None, device=0):
    """Batched form of :func:`isosurface`: element v of the result is
    ``isosurface(fields[v], methods[v], *axes[v])``, computed for all volumes
    together (one count/scan pass and one emit launch per kernel
    instantiation instead of one call each).

    ``methods`` is one algorithm object for all volumes or a list (default
    ``MarchingCubes()``); ``axes`` is ``None`` (``-1:1`` on every axis) or a
    list of ``(X, Y, Z)`` triples (``None`` entries allowed).  Fields are all
    host arrays or all CUDA tensors on one device; device fields give device
    results, which are views into one vertex buffer per output type and one
    face buffer."""
    n = len(fields)
    if methods is None or isinstance(methods, AbstractMeshingAlgorithm):
        methods = [methods or MarchingCubes()] * n
    if axes is None:
        axes = [None] * n
    if len(methods) != n or len(axes) != n:
        raise ValueError("fields, methods and axes must have the same length")
    for m in methods:
        if not isinstance(m, AbstractMeshingAlgorithm) or m.method < 0:
            raise TypeError(f"no isosurface method for {type(m).__name__}")
    if n == 0:
        return []
    prepared = [_field(f) for f in fields]
    on_dev = {d for _, _, d in prepared}
    if len(on_dev) != 1:
        raise ValueError("fields must be all host arrays or all CUDA tensors")
    on_dev = on_dev.pop()
    if on_dev and len({k.device for _, k, _ in prepared}) != 1:
        raise ValueError("device fields must be on one device")
    F_arr = (MshField * n)(*[f for f, _, _ in prepared])
    P_arr = (MshParams * n)(*[_params(m, *(ax if ax is not None else ((-1, 1),) * 3))
                             for m, ax in zip(methods, axes)])
    L = _capi.lib()
    kinds = []
    for v in range(n):
        ok = ctypes.c_int32()
        _capi.check_status(L.msh_output_kind(ctypes.byref(F_arr[v]), ctypes.byref(P_arr[v]),
                                             ctypes.byref(ok)), f"volume {v}")
        kinds.append(ok.value)
    dev = prepared[0][1].device if on_dev else None
    ctx = context(dev.index if on_dev else device)
    nv = (ctypes.c_int64 * n)()
    nf = (ctypes.c_int64 * n)()
    out = []
    with ctx.lock:
        if on_dev:
            torch = _torch()
            stream = torch.cuda.current_stream(dev)
            ctx.check(L.msh_set_stream(ctx.handle, ctypes.c_void_p(stream.cuda_stream)))
        else:
            ctx.check(L.msh_set_stream(ctx.handle, None))
        ctx.check(L.msh_batch_count(ctx.handle, n, F_arr, P_arr, nv, nf))
        vp = (ctypes.c_void_p * n)()
        fp = (ctypes.c_void_p * n)()
        if on_dev:
            tot32 = sum(nv[v] for v in range(n) if kinds[v] == MSH_FLOAT32)
            tot64 = sum(nv[v] for v in range(n) if kinds[v] != MSH_FLOAT32)
            V32 = torch.empty((tot32, 3), dtype=torch.float32, device=dev)
            V64 = torch.empty((tot64, 3), dtype=torch.float64, device=dev)
            FF = torch.empty((sum(nf), 3), dtype=torch.int64, device=dev)
            o32 = o64 = of = 0
            for v in range(n):
                if kinds[v] == MSH_FLOAT32:
                    V = V32[o32:o32 + nv[v]]
                    o32 += nv[v]
                else:
                    V = V64[o64:o64 + nv[v]]
                    o64 += nv[v]
                Fv = FF[of:of + nf[v]]
                of += nf[v]
                vp[v] = V.data_ptr() if nv[v] else None
                fp[v] = Fv.data_ptr() if nf[v] else None
                out.append((V, Fv))
            mem = MSH_DEVICE
        else:
            for v in range(n):
                V = np.empty((nv[v], 3), dtype=np.float32 if kinds[v] == MSH_FLOAT32 else np.float64)
                Fv = np.empty((nf[v], 3), dtype=np.int64)
                vp[v] = V.ctypes.data if nv[v] else None
                fp[v] = Fv.ctypes.data if nf[v] else None
                out.append((V, Fv))
            mem = MSH_HOST
        ctx.check(L.msh_batch_extract(ctx.handle, n, F_arr, P_arr, vp, fp, mem))
    return out


def autotune(sdf, method=None, *args, X=(-1, 1), Y=(-1, 1), Z=(-1, 1), device=0):
    """Time every emit kernel family on this field and method and keep the
    fastest for its (method, element type, size class) in the device's
    context (size classes by cell count: below 2^20, below 2^23, larger);
    later calls of isosurface, isosurface_batch and the server on that
    device use it for volumes of that class.  Returns the chosen family (3
    lanes along x with half-depth bricks, 2 lanes along x, 1 lanes along z,
    0 generic; see KERNEL_FAMILIES)."""
    method, X, Y, Z = _split_args(method, args, X, Y, Z)
    f, keep, on_dev = _field(sdf)
    p = _params(method, X, Y, Z)
    ctx = context(keep.device.index if on_dev else device)
    fam = ctypes.c_int32()
    with ctx.lock:
        if on_dev:
            ctx.check(_capi.lib().msh_set_stream(
                ctx.handle, ctypes.c_void_p(_torch().cuda.current_stream(keep.device).cuda_stream)))
        ctx.check(_capi.lib().msh_autotune(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                                           ctypes.byref(fam)))
    return fam.value


def autotune_batch(fields, methods=None, axes=None):
    """:func:`autotune` for a batch of device fields: each (method, element
    type, size class) group of the batch is timed as one batched emit with
    every kernel family, and the fastest is kept for that group; later
    batches (and single calls) on the device use it.  Returns the family
    chosen for each volume."""
    n = len(fields)
    if methods is None or isinstance(methods, AbstractMeshingAlgorithm):
        methods = [methods or MarchingCubes()] * n
    if axes is None:
        axes = [None] * n
    if n == 0:
        return []
    prepared = [_field(f) for f in fields]
    if not all(d for _, _, d in prepared) or len({k.device for _, k, _ in prepared}) != 1:
        raise ValueError("autotune_batch takes CUDA tensors on one device")
    dev = prepared[0][1].device
    F_arr = (MshField * n)(*[f for f, _, _ in prepared])
    P_arr = (MshParams * n)(*[_params(m, *(ax if ax is not None else ((-1, 1),) * 3))
                             for m, ax in zip(methods, axes)])
    chosen = (ctypes.c_int32 * n)()
    ctx = context(dev.index)
    with ctx.lock:
        ctx.check(_capi.lib().msh_set_stream(
            ctx.handle, ctypes.c_void_p(_torch().cuda.current_stream(dev).cuda_stream)))
        ctx.check(_capi.lib().msh_autotune_batch(ctx.handle, n, F_arr, P_arr, chosen))
    return list(chosen)


def set_kernel_family(method, dtype, family, device=0):
    """Choose the emit kernel family for (method, element type) on a device,
    for every size class (-1: default order).  Results do not depend on it."""
    m = method.method if isinstance(method, AbstractMeshingAlgorithm) else int(method)
    name = getattr(dtype, "__name__", None) or str(dtype).replace("torch.", "")
    kd = MSH_FLOAT32 if name == "float32" else MSH_FLOAT64
    ctx = context(device)
    with ctx.lock:
        ctx.check(_capi.lib().msh_set_kernel_family(ctx.handle, m, kd, int(family)))


FIELD_KINDS = {"sphere": 0, "torus": 1, "gyroid": 2, "value_noise": 3, "blobs": 4}


def generate_field(kind, shape, dtype="float32", seed=0, p0=None, p1=None, device=0):
    """Synthetic field generated on the device; returns an x-fastest CUDA
    tensor indexed [x, y, z].  Defaults: sphere r=0.5, torus R=0.6 r=0.2,
    gyroid on [-2pi, 2pi]^3."""
    torch = _torch()
    k = FIELD_KINDS[kind]
    defaults = {0: (0.5, 0.0), 1: (0.6, 0.2), 2: (2 * np.pi, 0.0), 3: (0.0, 0.0), 4: (0.0, 0.0)}
    d0, d1 = defaults[k]
    p0 = d0 if p0 is None else p0
    p1 = d1 if p1 is None else p1
    nx, ny, nz = shape
    tdt = torch.float32 if dtype in ("float32", np.float32, torch.float32) else torch.float64
    kd = MSH_FLOAT32 if tdt == torch.float32 else MSH_FLOAT64
    dev = torch.device("cuda", device)
    buf = torch.empty((nz, ny, nx), dtype=tdt, device=dev)
    ctx = context(device)
    with ctx.lock:
        ctx.check(_capi.lib().msh_set_stream(
            ctx.handle, ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)))
        ctx.check(_capi.lib().msh_generate_field(ctx.handle, k, kd, nx, ny, nz,
                                                 seed & 0xFFFFFFFFFFFFFFFF, float(p0), float(p1),
                                                 ctypes.c_void_p(buf.data_ptr())))
    return buf.permute(2, 1, 0)


def weld(vertices, faces, device=0):
    """Merge vertices whose three coordinates have identical bit patterns.

    Each group of equal vertices becomes its first appearance, the welded
    vertices keep first-appearance order, and the 1-based faces are
    renumbered to them; returns new ``(vertices, faces)`` of the same kind
    as the arguments (CUDA tensors, or numpy arrays, which travel to the
    device and back).  Equality is bitwise: -0.0 and 0.0 stay apart, NaNs
    weld only with the same payload.  The opt-in step after Marching Cubes,
    whose cubes repeat the vertices they share (src/algorithmtypes.jl:17-19).
    Two cubes can interpolate a shared edge from opposite ends
    (src/lut/mc.jl _mc_edge_list), and then its two copies may differ in the
    last bit and stay apart."""
    torch = _torch()
    host = isinstance(vertices, np.ndarray)
    if host:
        dev = torch.device("cuda", device)
        V = torch.from_numpy(np.ascontiguousarray(vertices)).to(dev)
        F = torch.from_numpy(np.ascontiguousarray(faces, dtype=np.int64)).to(dev)
    else:
        if not (_is_device_tensor(vertices) and _is_device_tensor(faces)):
            raise TypeError("weld takes CUDA tensors or numpy arrays")
        V = vertices.contiguous()
        F = faces.to(torch.int64).contiguous().clone()  # renumbered in place
    if V.ndim != 2 or V.shape[1] != 3 or F.ndim != 2 or F.shape[1] != 3:
        raise ValueError("vertices and faces must be (n, 3) arrays")
    if V.dtype not in (torch.float32, torch.float64):
        raise TypeError("vertex coordinates must be Float32 or Float64")
    kind = MSH_FLOAT32 if V.dtype == torch.float32 else MSH_FLOAT64
    out = torch.empty_like(V)
    n = ctypes.c_int64()
    ctx = context(V.device.index)
    L = _capi.lib()
    with ctx.lock:
        ctx.check(L.msh_set_stream(
            ctx.handle, ctypes.c_void_p(torch.cuda.current_stream(V.device).cuda_stream)))
        ctx.check(L.msh_weld(ctx.handle, kind, V.shape[0],
                             ctypes.c_void_p(V.data_ptr() if V.numel() else None), F.shape[0],
                             ctypes.c_void_p(F.data_ptr() if F.numel() else None),
                             ctypes.c_void_p(out.data_ptr() if out.numel() else None),
                             ctypes.byref(n)))
    out = out[:n.value].clone()
    if host:
        return out.cpu().numpy(), F.cpu().numpy()
    return out, F


from .io import NrrdInfo, read_nrrd, read_ply, write_nrrd, write_ply  # noqa: E402
from .serve import FieldSpec, Server  # noqa: E402
