"""GPU parity of Marching Cubes: the HIP library (through the C ABI) against
the CPU oracle on the same field bits.  Faces must be identical and vertex
coordinates bit-identical (the library is built with -ffp-contract=off)."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, noise, sphere_field

pytestmark = pytest.mark.gpu


def run_both(pkg, oracle, f, iso, X=(-1, 1), Y=(-1, 1), Z=(-1, 1), host=False):
    meth = pkg.MarchingCubes(iso=iso)
    src = f if host else dev_field(f)
    V, F = pkg.isosurface(src, meth, X, Y, Z)
    if not host:
        V, F = V.cpu().numpy(), F.cpu().numpy()
    Vo, Fo = oracle.isosurface(f, "mc", iso=iso, X=X, Y=Y, Z=Z)
    assert_same_mesh(V, F, Vo, Fo)
    return V, F


def test_sphere_128_float32_int_iso(pkg, oracle):
    # 128^3 Float32 sphere sqrt(x^2+y^2+z^2)-0.5 on [-1,1]^3, MarchingCubes(iso=0)
    f = sphere_field(128, 0.5, np.float32)
    V, F = run_both(pkg, oracle, f, 0)
    assert V.dtype == np.float32 and len(F) > 10000


@pytest.mark.parametrize("iso", [0.0, np.float32(0.0), 0, 0.013])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_types(pkg, oracle, iso, dtype):
    f = sphere_field(47, 0.55, dtype)
    run_both(pkg, oracle, f, iso)


@pytest.mark.parametrize("axes", [
    ((-1.0, 1.0), (-1.0, 1.0), (-1.0, 1.0)),
    ((np.float32(-1.5), np.float32(2.25)),) * 3,
    ((0, 10), (np.float32(-0.3), np.float32(0.7)), (-2.5, 3.0)),
    ((3, -3), (1, 0), (0.1, 0.2)),
])
def test_axis_kinds(pkg, oracle, axes):
    f = noise((33, 21, 70), 7, np.float32)
    run_both(pkg, oracle, f, 0, *axes)


@pytest.mark.parametrize("shape", [(2, 2, 2), (2, 40, 3), (65, 3, 130), (37, 29, 70),
                                   (31, 5, 65), (32, 6, 64), (63, 9, 129), (200, 4, 5),
                                   (5, 200, 4), (4, 5, 300)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ragged_shapes(pkg, oracle, shape, dtype):
    f = noise(shape, sum(shape), dtype)
    run_both(pkg, oracle, f, 0.05)


def test_host_arrays(pkg, oracle):
    f = noise((20, 30, 40), 1, np.float64)
    V, F = run_both(pkg, oracle, f, 0.0, host=True)
    assert isinstance(V, np.ndarray)


def test_nonfinite_values(pkg, oracle):
    f = noise((24, 25, 26), 2, np.float32)
    f[3, 4, 5] = np.nan
    f[10, :, 7] = np.inf
    f[:, 12, 12] = -np.inf
    f[20, 20, :] = np.nan
    run_both(pkg, oracle, f, 0)
    run_both(pkg, oracle, f.astype(np.float64), 0.0)


def test_no_surface_and_degenerate_sizes(pkg, oracle):
    V, F = run_both(pkg, oracle, np.ones((30, 30, 30), np.float32), 0)
    assert len(V) == 0 and len(F) == 0
    for shape in ((0, 4, 4), (4, 4, 0)):
        V, F = pkg.isosurface(dev_field(np.zeros(shape, np.float32)), pkg.MarchingCubes())
        assert V.shape == (0, 3) and F.shape == (0, 3)
    with pytest.raises(ValueError, match="endpoints differ"):
        pkg.isosurface(dev_field(np.zeros((4, 1, 4), np.float32)), pkg.MarchingCubes())
    # length 1 with equal endpoints is allowed and has no cells
    V, F = pkg.isosurface(dev_field(np.zeros((4, 1, 4), np.float32)), pkg.MarchingCubes(),
                          (-1, 1), (0, 0), (-1, 1))
    assert len(F) == 0


def test_default_method_is_mc(pkg, oracle):
    # test/runtests.jl:76-79 and src/isosurface.jl:30-32
    f = noise((10, 10, 10), 5, np.float64)
    V1, F1 = pkg.isosurface(dev_field(f))
    V2, F2 = pkg.isosurface(dev_field(f), pkg.MarchingCubes())
    assert np.array_equal(V1.cpu().numpy(), V2.cpu().numpy())
    assert np.array_equal(F1.cpu().numpy(), F2.cpu().numpy())
    V3, F3 = pkg.isosurface(dev_field(f), (-1, 1), (-1, 1), (-1, 1))
    assert np.array_equal(F1.cpu().numpy(), F3.cpu().numpy())


def test_repeated_calls_grow_and_shrink(pkg, oracle):
    for shape in ((40, 40, 40), (90, 20, 70), (12, 12, 12), (100, 100, 30)):
        run_both(pkg, oracle, noise(shape, shape[0], np.float32), 0)


def test_value_noise_256(pkg, oracle):
    f = pkg.generate_field("value_noise", (256, 256, 256), "float32", seed=3)
    run_both(pkg, oracle, f.cpu().numpy(), 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_dense_crossings(pkg, oracle, dtype):
    # white noise: nearly every cell crosses, so chunks overflow the emit
    # queue and are taken one z step at a time
    f = np.random.default_rng(11).standard_normal((70, 40, 90)).astype(dtype)
    run_both(pkg, oracle, f, 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_long_z_rows(pkg, oracle, dtype):
    # long z rows, split into z segments (units) by the host
    f = noise((9, 7, 1500), 4, dtype)
    run_both(pkg, oracle, f, 0.02)
