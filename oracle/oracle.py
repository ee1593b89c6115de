"""ctypes front end of the CPU oracle (liboracle.so, built from cpu_ref.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker, never as the product.

Argument conventions mirror the reference's ``isosurface(A, method, X, Y, Z)``
(src/isosurface.jl:30-32, src/marching_cubes.jl:27, src/marching_tetrahedra.jl:129):
``sdf`` is indexed ``sdf[x, y, z]`` with x fastest in memory (Julia's
column-major order), the result is ``(vertices (V,3), faces (F,3) int64,
1-based)``.  Julia scalar types are inferred from the Python values:
``int`` -> Int64, ``float`` -> Float64, ``numpy.float32`` -> Float32.
"""
import ctypes
import os
from pathlib import Path

import numpy as np

K_INT64, K_F32, K_F64 = 0, 1, 2
METHODS = {"mc": 0, "mt": 1, "sn": 2}
_HERE = Path(__file__).resolve().parent
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        # MSH_ORACLE_LIB names another build of cpu_ref.cpp (the sanitizer
        # build, `make -C oracle sanitize`); test infrastructure only
        path = Path(os.environ.get("MSH_ORACLE_LIB", _HERE / "liboracle.so"))
        if not path.exists():
            raise RuntimeError(f"{path} is not built: run `make -C {_HERE}`")
        L = ctypes.CDLL(str(path))
        P, I64, D, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int
        L.oracle_isosurface.argtypes = [I, I, P, I64, I64, I64, I, D, I64, I, D, I64, P, P, P,
                                        ctypes.POINTER(P), ctypes.POINTER(I64),
                                        ctypes.POINTER(P), ctypes.POINTER(I64),
                                        ctypes.POINTER(I)]
        L.oracle_isosurface.restype = I
        L.oracle_isosurface_xslab.argtypes = [I, P, I64, I64, I64, I64, I64, I64, I64, I, D,
                                              I64, P, P, P, ctypes.POINTER(P),
                                              ctypes.POINTER(I64), ctypes.POINTER(P),
                                              ctypes.POINTER(I64), ctypes.POINTER(I)]
        L.oracle_isosurface_xslab.restype = I
        L.oracle_free.argtypes = [P]
        L.oracle_set_threads.argtypes = [I]
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_linrange.argtypes = [I, D, D, I64, P]
        L.oracle_linrange.restype = I
        L.oracle_cubeindex_f64.argtypes = [P, D]
        L.oracle_cubeindex_f64.restype = I
        L.oracle_vertex_interp_i64.argtypes = [I64, P, P, I64, I64, P]
        _LIB = L
    return _LIB


def kind_of(v):
    """Julia type of a Python/numpy scalar (Int64, Float32 or Float64)."""
    if isinstance(v, (bool, np.bool_)):
        raise TypeError("Bool is not supported")
    if isinstance(v, (int, np.integer)):
        return K_INT64
    if isinstance(v, np.float32):
        return K_F32
    if isinstance(v, (float, np.float64)):
        return K_F64
    raise TypeError(f"unsupported scalar type {type(v)}")


def _endpoints(R):
    """first(R), last(R) of an axis argument: a (first, last) pair, a Python
    range or a 1-D numpy array."""
    if isinstance(R, range):
        return R[0], R[-1]
    if isinstance(R, np.ndarray):
        return R[0], R[-1]
    a, b = R
    return a, b


def isosurface(sdf, method="mc", iso=0.0, eps=1e-3, X=(-1, 1), Y=(-1, 1), Z=(-1, 1),
               threads=1):
    """threads > 1 runs Marching Cubes over x-slabs on that many threads (a
    CPU baseline over all cores; same result as the single-threaded form)."""
    sdf = np.asarray(sdf)
    if sdf.ndim != 3:
        raise ValueError("sdf must be 3-D")
    if sdf.dtype == np.float32:
        dtype = K_F32
    elif sdf.dtype == np.float64:
        dtype = K_F64
    else:
        raise TypeError(f"unsupported field dtype {sdf.dtype}")
    a = np.asfortranarray(sdf)
    nx, ny, nz = a.shape
    kinds, lo, hi = [], [], []
    for R in (X, Y, Z):
        f, l = _endpoints(R)
        k = kind_of(f)
        kinds.append(k)
        lo.append(float(f))
        hi.append(float(l))
    kinds = (ctypes.c_int * 3)(*kinds)
    lo = (ctypes.c_double * 3)(*lo)
    hi = (ctypes.c_double * 3)(*hi)
    ik, ek = kind_of(iso), kind_of(eps)
    iso_i = int(iso) if ik == K_INT64 else 0
    eps_i = int(eps) if ek == K_INT64 else 0
    vp, fp = ctypes.c_void_p(), ctypes.c_void_p()
    nv, nf, ok = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    L = lib()
    L.oracle_set_threads(int(threads))
    rc = L.oracle_isosurface(METHODS[method] if isinstance(method, str) else int(method), dtype,
                             a.ctypes.data, nx, ny, nz, ik, float(iso), iso_i, ek, float(eps),
                             eps_i, kinds, lo, hi, ctypes.byref(vp), ctypes.byref(nv),
                             ctypes.byref(fp), ctypes.byref(nf), ctypes.byref(ok))
    if rc != 0:
        raise ValueError(L.oracle_last_error().decode())
    odt = np.float32 if ok.value == K_F32 else np.float64
    try:
        V = np.ctypeslib.as_array((ctypes.c_char * (nv.value * 3 * np.dtype(odt).itemsize))
                                  .from_address(vp.value)).view(odt).reshape(-1, 3).copy() \
            if nv.value else np.zeros((0, 3), odt)
        F = np.ctypeslib.as_array((ctypes.c_int64 * (nf.value * 3)).from_address(fp.value)) \
            .reshape(-1, 3).copy() if nf.value else np.zeros((0, 3), np.int64)
    finally:
        L.oracle_free(vp)
        L.oracle_free(fp)
    return V, F


def _collect(vp, nv, fp, nf, ok):
    odt = np.float32 if ok.value == K_F32 else np.float64
    L = lib()
    try:
        V = np.ctypeslib.as_array((ctypes.c_char * (nv.value * 3 * np.dtype(odt).itemsize))
                                  .from_address(vp.value)).view(odt).reshape(-1, 3).copy() \
            if nv.value else np.zeros((0, 3), odt)
        F = np.ctypeslib.as_array((ctypes.c_int64 * (nf.value * 3)).from_address(fp.value)) \
            .reshape(-1, 3).copy() if nf.value else np.zeros((0, 3), np.int64)
    finally:
        L.oracle_free(vp)
        L.oracle_free(fp)
    return V, F


def isosurface_xslab(slab, nx, xoff, xlo, xhi, iso=0.0, X=(-1, 1), Y=(-1, 1), Z=(-1, 1)):
    """Marching Cubes over the cells with x index xlo..xhi (1-based) of an
    array that is nx samples long along x, given only its x-planes
    xoff+1 .. xoff+slab.shape[0] (``slab`` indexed [x, y, z]).  Returns the
    slice of the whole array's mesh that these cells emit, with faces
    counted from the slice's first vertex: the reference's loop runs x
    outermost (src/marching_cubes.jl:40) and a cube's faces name only its
    own vertices, so the whole mesh is these slices concatenated with
    shifted faces."""
    slab = np.asarray(slab)
    if slab.dtype == np.float32:
        dtype = K_F32
    elif slab.dtype == np.float64:
        dtype = K_F64
    else:
        raise TypeError(f"unsupported field dtype {slab.dtype}")
    a = np.asfortranarray(slab)
    sx, ny, nz = a.shape
    kinds, lo, hi = [], [], []
    for R in (X, Y, Z):
        f, l = _endpoints(R)
        kinds.append(kind_of(f))
        lo.append(float(f))
        hi.append(float(l))
    ik = kind_of(iso)
    vp, fp = ctypes.c_void_p(), ctypes.c_void_p()
    nv, nf, ok = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
    L = lib()
    rc = L.oracle_isosurface_xslab(dtype, a.ctypes.data, nx, ny, nz, xoff, sx, xlo, xhi, ik,
                                   float(iso), int(iso) if ik == K_INT64 else 0,
                                   (ctypes.c_int * 3)(*kinds), (ctypes.c_double * 3)(*lo),
                                   (ctypes.c_double * 3)(*hi), ctypes.byref(vp), ctypes.byref(nv),
                                   ctypes.byref(fp), ctypes.byref(nf), ctypes.byref(ok))
    if rc != 0:
        raise ValueError(L.oracle_last_error().decode())
    return _collect(vp, nv, fp, nf, ok)


def linrange(first, last, n):
    out = np.zeros(max(n, 0), np.float64)
    rc = lib().oracle_linrange(kind_of(first), float(first), float(last), n, out.ctypes.data)
    if rc != 0:
        raise ValueError(lib().oracle_last_error().decode())
    return out


def cubeindex(vals, iso):
    v = np.ascontiguousarray(vals, np.float64)
    return lib().oracle_cubeindex_f64(v.ctypes.data, float(iso))


def vertex_interp_int(iso, p1, p2, v1, v2):
    a = np.asarray(p1, np.int64)
    b = np.asarray(p2, np.int64)
    out = np.zeros(3)
    lib().oracle_vertex_interp_i64(iso, a.ctypes.data, b.ctypes.data, v1, v2, out.ctypes.data)
    return tuple(out)


def weld(V, F):
    """Exact-coordinate vertex welding, the definition the GPU's msh_weld
    follows (the reference has no welding function; its Marching Cubes
    repeats shared vertices, src/algorithmtypes.jl:17-19): vertices with
    bit-identical (x, y, z) merge into their first appearance, welded
    vertices keep first-appearance order, faces (1-based) are renumbered.
    numpy's sort-based ``unique`` over the raw bytes of each vertex; the
    tests pin it against a dictionary walk and against the defining
    properties."""
    V = np.ascontiguousarray(V)
    F = np.ascontiguousarray(F, dtype=np.int64)
    n = V.shape[0]
    if F.size and (F.min() < 1 or F.max() > n):
        raise ValueError("face index outside 1..nverts")
    if n == 0:
        return V.copy(), F.copy()
    key = V.view(np.dtype((np.void, V.dtype.itemsize * 3))).ravel()
    _, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")      # unique keys by first appearance
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    newid = rank[inv.ravel()]
    return V[first[order]].copy(), newid[F - 1] + 1
