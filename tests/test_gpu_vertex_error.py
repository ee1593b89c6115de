"""The mesh of a Float32 field against the mesh of the same samples as
Float64: the same faces (classification compares exact values), and
vertex coordinates within a few Float32 ulps of the unit range (the only
difference is the arithmetic type of the interpolation weight)."""
import numpy as np
import pytest

from helpers import dev_field, sphere_field

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["mc", "mt", "sn"])
def test_float32_field_against_float64(pkg, method):
    import torch
    m = {"mc": pkg.MarchingCubes(iso=0), "mt": pkg.MarchingTetrahedra(iso=0),
         "sn": pkg.NaiveSurfaceNets(iso=0)}[method]
    f = sphere_field(96, 0.55, np.float32)
    V32, F32 = pkg.isosurface(dev_field(f), m)
    V64, F64 = pkg.isosurface(dev_field(f.astype(np.float64)), m)
    assert torch.equal(F32, F64) and V32.shape == V64.shape
    d = (V32.double() - V64).abs().max().item()
    # coordinates lie in [-1, 1]: 4 Float32 ulps of 1.0
    assert d <= 4 * 2.0 ** -23, d
