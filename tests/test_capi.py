"""CPU checks of the C ABI: the library loads, exports every function
include/meshing_hip.h declares, and resolves Julia's type promotion into the
same output type as the oracle -- all without touching a GPU."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "meshing_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(msh_\w+)\s*\(", text, re.M)))


def test_header_lists_expected_entry_points():
    names = declared_functions()
    assert "msh_count" in names and "msh_extract" in names and "msh_create" in names
    assert len(names) >= 9


def test_library_exports_every_declared_symbol(pkg):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    lib = capi.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(capi.EXPORTS) == set(declared_functions())


def test_struct_layouts(pkg):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    assert ctypes.sizeof(capi.MshField) == 40
    assert capi.MshParams.iso.offset == 24
    assert ctypes.sizeof(capi.MshParams) == 24 + 2 * 8 + 2 * 8 + 6 * 8


def out_kind(pkg, dtype, method, X=(-1, 1)):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    f = capi.MshField(capi.MSH_FLOAT32 if dtype == np.float32 else capi.MSH_FLOAT64,
                      capi.MSH_HOST, 4, 4, 4, None)
    p = pkg._params(method, X, (-1, 1), (-1, 1))
    k = ctypes.c_int32()
    rc = capi.lib().msh_output_kind(ctypes.byref(f), ctypes.byref(p), ctypes.byref(k))
    assert rc == 0
    return np.float32 if k.value == capi.MSH_FLOAT32 else np.float64


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("iso", [0, 0.0, np.float32(0.25)])
@pytest.mark.parametrize("X", [(-1, 1), (-1.0, 1.0), (np.float32(-1), np.float32(1))])
def test_output_kind_matches_oracle(pkg, oracle, dtype, iso, X):
    f = np.random.default_rng(0).standard_normal((4, 4, 4)).astype(dtype)
    for method, name in ((pkg.MarchingCubes(iso=iso), "mc"),
                         (pkg.NaiveSurfaceNets(iso=iso), "sn"),
                         (pkg.MarchingTetrahedra(iso=iso), "mt"),
                         (pkg.MarchingTetrahedra(iso=iso, eps=np.float32(1e-3)), "mt")):
        V, _ = oracle.isosurface(f, name, iso=iso, eps=getattr(method, "eps", 1e-3), X=X)
        assert out_kind(pkg, dtype, method, X) == V.dtype, (name, method)


def test_unsupported_types_raise(pkg):
    with pytest.raises(TypeError):
        pkg._params(pkg.MarchingCubes(iso="0"), (-1, 1), (-1, 1), (-1, 1))
    with pytest.raises(TypeError):
        pkg.isosurface(np.zeros((3, 3, 3), dtype=np.int64))
    with pytest.raises(TypeError):
        pkg.isosurface(np.zeros((3, 3, 3), dtype=np.float16))
    with pytest.raises(ValueError):
        pkg.isosurface(np.zeros((3, 3)))
