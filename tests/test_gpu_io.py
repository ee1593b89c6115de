"""The reference example's flow on the GPU path (examples/nrrd.jl:8-21):
an Int16 CT-like NRRD volume read as Float32, extracted on the device with
MarchingTetrahedra(iso=100), welded and saved as PLY; every step checked
against the oracle."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, sphere_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["mt", "mc"])
def test_nrrd_extract_ply(pkg, oracle, tmp_path, method):
    ct = np.round(sphere_field(40, 0.5, np.float64) * 1000).astype(np.int16)
    pkg.write_nrrd(tmp_path / "ct.nrrd", ct, encoding="gzip", spacings=(0.5, 0.5, 1.0))
    a, info = pkg.read_nrrd(tmp_path / "ct.nrrd", dtype=np.float32)
    X, Y, Z = info.axes()
    m = (pkg.MarchingTetrahedra(iso=np.float32(100)) if method == "mt"
         else pkg.MarchingCubes(iso=np.float32(100)))
    V, F = pkg.isosurface(dev_field(a), m, X, Y, Z)
    Vo, Fo = oracle.isosurface(a, method, iso=np.float32(100), X=X, Y=Y, Z=Z)
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
    Vw, Fw = pkg.weld(V, F)
    pkg.write_ply(tmp_path / "ct.ply", Vw, Fw)
    V2, F2 = pkg.read_ply(tmp_path / "ct.ply")
    Vwo, Fwo = oracle.weld(Vo, Fo)
    assert_same_mesh(V2, F2, Vwo, Fwo)
