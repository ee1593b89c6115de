"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py) against the CPU oracle and, on the GPU,
against the HIP library through the C ABI: faces identical, vertex
coordinates bit-identical."""
import json
from pathlib import Path

import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))
TYPES = {"int64": int, "float32": np.float32, "float64": float}


def load(path):
    with np.load(path, allow_pickle=False) as z:
        field, V, F = z["field"], z["V"], z["F"]
        meta = json.loads(str(z["meta"]))
    val = lambda p: TYPES[p[0]](p[1])  # noqa: E731
    iso = val(meta["iso"])
    axes = [tuple(val(p) for p in meta[a]) for a in "XYZ"]
    return field, V, F, meta["method"], iso, meta["eps"], axes


def method_obj(pkg, method, iso, eps):
    if method == "mc":
        return pkg.MarchingCubes(iso=iso)
    if method == "mt":
        return pkg.MarchingTetrahedra(iso=iso, eps=eps)
    return pkg.NaiveSurfaceNets(iso=iso)


def test_fixtures_present():
    names = {p.stem.split("_")[0] for p in GOLDEN}
    assert names == {"mc", "mt", "sn"}, names


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.stem)
def test_oracle_reproduces_fixture(oracle, path):
    field, V, F, method, iso, eps, (X, Y, Z) = load(path)
    kw = dict(iso=iso, X=X, Y=Y, Z=Z)
    if eps is not None:
        kw["eps"] = eps
    Vo, Fo = oracle.isosurface(field, method, **kw)
    assert_same_mesh(Vo, Fo, V, F)


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: p.stem)
@pytest.mark.parametrize("where", ["device", "host"])
def test_library_reproduces_fixture(pkg, path, where):
    field, V, F, method, iso, eps, (X, Y, Z) = load(path)
    src = dev_field(field) if where == "device" else field
    Vg, Fg = pkg.isosurface(src, method_obj(pkg, method, iso, eps), X, Y, Z)
    if where == "device":
        Vg, Fg = Vg.cpu().numpy(), Fg.cpu().numpy()
    assert_same_mesh(Vg, Fg, V, F)
