"""ctypes binding of libmeshing_hip.so (C ABI declared in include/meshing_hip.h).

There is no fallback: if the library is missing or fails to load, every call
raises.  The structures mirror ``msh_field`` and ``msh_params`` field for
field.
"""
import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmeshing_hip.so"

MSH_OK, MSH_ERR_ARGUMENT, MSH_ERR_OUT_OF_MEMORY, MSH_ERR_DEVICE, MSH_ERR_STATE, \
    MSH_ERR_UNSUPPORTED = range(6)
MSH_MARCHING_CUBES, MSH_MARCHING_TETRAHEDRA, MSH_NAIVE_SURFACE_NETS = range(3)
MSH_INT64, MSH_FLOAT32, MSH_FLOAT64 = range(3)
MSH_HOST, MSH_DEVICE = range(2)

# every symbol the header declares; checked by the CPU test suite
EXPORTS = ("msh_create", "msh_destroy", "msh_set_stream", "msh_synchronize",
           "msh_last_error", "msh_output_kind", "msh_count", "msh_extract",
           "msh_batch_count", "msh_batch_extract", "msh_autotune", "msh_set_kernel_family",
           "msh_generate_field", "msh_weld", "msh_isosurface_noalloc",
           "msh_batch_graph_launches", "msh_autotune_batch")


class MshField(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("memory", ctypes.c_int32),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("data", ctypes.c_void_p)]


class MshParams(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("iso_kind", ctypes.c_int32),
                ("eps_kind", ctypes.c_int32), ("axis_kind", ctypes.c_int32 * 3),
                ("iso", ctypes.c_double), ("eps", ctypes.c_double),
                ("iso_int", ctypes.c_int64), ("eps_int", ctypes.c_int64),
                ("first", ctypes.c_double * 3), ("last", ctypes.c_double * 3)]


class MeshingError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check_status(rc, what):
    """Raise for a status returned by a call that has no context (no message)."""
    if rc == MSH_OK:
        return
    if rc == MSH_ERR_ARGUMENT:
        raise ValueError(f"{what}: invalid argument")
    if rc == MSH_ERR_UNSUPPORTED:
        raise TypeError(f"{what}: unsupported element type")
    raise MeshingError(rc, f"{what}: status {rc}")


_lib = None
_lib_lock = threading.Lock()


def lib():
    """Load libmeshing_hip.so once; raise if it is not built."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `make -C {LIB_PATH.parent}` "
                    "(hipcc, gfx950).  There is no CPU fallback.")
            # torch ships its own HIP runtime under the same soname; load it
            # first so that the library binds to it and the process has one
            # HIP runtime whichever side touches the GPU first
            try:
                import torch  # noqa: F401
            except ImportError:
                pass
            L = ctypes.CDLL(str(LIB_PATH))
            vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
            pf, pp = ctypes.POINTER(MshField), ctypes.POINTER(MshParams)
            L.msh_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
            L.msh_destroy.argtypes = [vp]
            L.msh_set_stream.argtypes = [vp, vp]
            L.msh_synchronize.argtypes = [vp]
            L.msh_last_error.argtypes = [vp]
            L.msh_last_error.restype = ctypes.c_char_p
            L.msh_output_kind.argtypes = [pf, pp, ctypes.POINTER(i32)]
            L.msh_count.argtypes = [vp, pf, pp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
            L.msh_extract.argtypes = [vp, pf, pp, vp, vp, i32]
            L.msh_batch_count.argtypes = [vp, i32, pf, pp, ctypes.POINTER(i64),
                                          ctypes.POINTER(i64)]
            L.msh_batch_extract.argtypes = [vp, i32, pf, pp, ctypes.POINTER(vp),
                                            ctypes.POINTER(vp), i32]
            L.msh_autotune.argtypes = [vp, pf, pp, ctypes.POINTER(i32)]
            L.msh_set_kernel_family.argtypes = [vp, i32, i32, i32]
            L.msh_generate_field.argtypes = [vp, i32, i32, i64, i64, i64, ctypes.c_uint64,
                                             ctypes.c_double, ctypes.c_double, vp]
            L.msh_weld.argtypes = [vp, i32, i64, vp, i64, vp, vp, ctypes.POINTER(i64)]
            L.msh_isosurface_noalloc.argtypes = [vp, pf, pp, vp, i64, vp, i64, vp]
            L.msh_batch_graph_launches.argtypes = [vp, ctypes.POINTER(i64), ctypes.POINTER(i64)]
            L.msh_autotune_batch.argtypes = [vp, i32, pf, pp, ctypes.POINTER(i32)]
            for name in EXPORTS:
                f = getattr(L, name)
                if name != "msh_last_error":
                    f.restype = ctypes.c_int
            _lib = L
        return _lib


class Context:
    """One msh context (one HIP stream, grow-only workspaces) per device."""

    def __init__(self, device=0):
        self.device = device
        self.handle = ctypes.c_void_p()
        self.lock = threading.Lock()
        rc = lib().msh_create(device, ctypes.byref(self.handle))
        if rc != MSH_OK:
            raise MeshingError(rc, f"msh_create(device={device}) failed with status {rc}")

    def check(self, rc):
        if rc != MSH_OK:
            msg = lib().msh_last_error(self.handle).decode()
            if rc == MSH_ERR_ARGUMENT:
                raise ValueError(msg)
            if rc == MSH_ERR_UNSUPPORTED:
                raise TypeError(msg)
            if rc == MSH_ERR_OUT_OF_MEMORY:
                raise MemoryError(msg)
            raise MeshingError(rc, msg)

    def close(self):
        if self.handle:
            lib().msh_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts = {}
_ctx_lock = threading.Lock()


def context(device=0):
    with _ctx_lock:
        c = _contexts.get(device)
        if c is None:
            c = _contexts[device] = Context(device)
        return c
