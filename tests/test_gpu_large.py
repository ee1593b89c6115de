"""Fields whose z-planes are too large for the 32-bit sample offsets of the
emit kernels' LDS bricks: the library falls back to the next kernel family
(in the end the generic emit, which addresses the field with 64-bit
offsets) and the mesh stays bit-identical to the oracle's.  The surface is
a thin cylinder, so the output (and the oracle's run time) stays small."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field

pytestmark = pytest.mark.gpu


def cylinder(nx, ny, nz, dtype):
    x = np.linspace(-1.0, 1.0, nx)[:, None, None]
    y = np.linspace(-1.0, 1.0, ny)[None, :, None]
    z = np.linspace(-1.0, 1.0, nz)[None, None, :]
    return (np.sqrt(x * x + y * y) - 0.5 + 0.01 * z).astype(dtype)


def check(pkg, oracle, f, name):
    m = {"mc": pkg.MarchingCubes(iso=0), "mt": pkg.MarchingTetrahedra(iso=0),
         "sn": pkg.NaiveSurfaceNets(iso=0)}[name]
    V, F = pkg.isosurface(dev_field(f), m)
    Vo, Fo = oracle.isosurface(f, name, iso=0)
    assert len(Fo) > 1000
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)


@pytest.mark.parametrize("name,shape,dtype", [
    # 16-17 Float32 planes of 6000 x 6000 exceed 2 GiB: generic Marching Cubes emit
    ("mc", (6000, 6000, 3), np.float32),
    ("mc", (6000, 6000, 2), np.float64),
    # 9-10 Float32 planes of 8000 x 8000 exceed 2 GiB: generic MT / SN emit
    ("mt", (8000, 8000, 2), np.float32),
    ("sn", (8000, 8000, 2), np.float32),
    ("mt", (7400, 7400, 2), np.float64),
])
def test_planes_past_32bit_brick_offsets(pkg, oracle, name, shape, dtype):
    check(pkg, oracle, cylinder(*shape, dtype), name)


def test_large_plane_in_a_batch(pkg, oracle):
    # a volume that needs the generic kernel next to small ones in one batch
    fields = [cylinder(6000, 6000, 3, np.float32), cylinder(40, 30, 20, np.float32),
              cylinder(33, 47, 25, np.float32)]
    res = pkg.isosurface_batch([dev_field(f) for f in fields], pkg.MarchingCubes(iso=0))
    for f, (V, F) in zip(fields, res):
        Vo, Fo = oracle.isosurface(f, "mc", iso=0)
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
