"""GPU parity of Marching Tetrahedra and Naive Surface Nets against the CPU
oracle (same field bits; faces identical, vertices bit-identical)."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, noise, sphere_field

pytestmark = pytest.mark.gpu


def run_both(pkg, oracle, f, name, iso, eps=1e-3, axes=((-1, 1), (-1, 1), (-1, 1)), host=False):
    meth = (pkg.MarchingTetrahedra(iso=iso, eps=eps) if name == "mt"
            else pkg.NaiveSurfaceNets(iso=iso))
    V, F = pkg.isosurface(f if host else dev_field(f), meth, *axes)
    if not host:
        V, F = V.cpu().numpy(), F.cpu().numpy()
    Vo, Fo = oracle.isosurface(f, name, iso=iso, eps=eps, X=axes[0], Y=axes[1], Z=axes[2])
    assert_same_mesh(V, F, Vo, Fo)
    return V, F


@pytest.mark.parametrize("name", ["mt", "sn"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_sphere(pkg, oracle, name, dtype):
    run_both(pkg, oracle, sphere_field(70, 0.6, dtype), name, 0)


@pytest.mark.parametrize("name", ["mt", "sn"])
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 3), (2, 40, 3), (65, 3, 130), (37, 29, 70),
                                   (31, 5, 65), (32, 6, 64), (17, 2, 129), (200, 4, 5),
                                   (5, 200, 4), (4, 5, 300)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ragged_shapes(pkg, oracle, name, shape, dtype):
    run_both(pkg, oracle, noise(shape, sum(shape) + 1, dtype), name, 0.02)


@pytest.mark.parametrize("iso,eps", [(0, 1e-3), (0.0, 1e-3), (np.float32(0.01), np.float32(1e-3)),
                                     (0, np.float32(0.01)), (0, 0), (0.05, 0.25)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mt_types(pkg, oracle, iso, eps, dtype):
    run_both(pkg, oracle, noise((23, 19, 41), 11, dtype), "mt", iso, eps)


@pytest.mark.parametrize("name", ["mt", "sn"])
@pytest.mark.parametrize("axes", [((0, 1), (0, 1), (0, 1)),
                                  ((np.float32(-2), np.float32(2)),) * 3,
                                  ((0, 10), (np.float32(-0.3), np.float32(0.7)), (-2.5, 3.0))])
def test_axis_kinds(pkg, oracle, name, axes):
    run_both(pkg, oracle, noise((30, 22, 67), 3, np.float32), name, 0, axes=axes)


@pytest.mark.parametrize("name", ["mt", "sn"])
def test_nonfinite_values(pkg, oracle, name):
    f = noise((24, 25, 26), 2, np.float32)
    f[3, 4, 5] = np.nan
    f[10, :, 7] = np.inf
    f[:, 12, 12] = -np.inf
    run_both(pkg, oracle, f, name, 0)


def test_mt_noisy_sphere_shape(pkg, oracle):
    # the shape of test/runtests.jl:152-172 (21^3, distance + noise,
    # iso = N - 2 sigma); the reference's MersenneTwister stream is not
    # reproducible here, so the noise is numpy's
    N = 10
    i = np.arange(-N, N + 1, dtype=np.float32)
    d = np.sqrt(i[:, None, None] ** 2 + i[None, :, None] ** 2 + i[None, None, :] ** 2)
    f = d.astype(np.float32) + 1.0 * np.random.default_rng(0).random((21, 21, 21))
    run_both(pkg, oracle, f, "mt", float(N - 2))


@pytest.mark.parametrize("name", ["mt", "sn"])
def test_zsegments_and_big(pkg, oracle, name):
    # long z rows are cut into segments; many x/y tiles
    run_both(pkg, oracle, noise((40, 35, 700), 5, np.float32), name, 0)
    f = pkg.generate_field("gyroid", (128, 128, 128), "float32")
    run_both(pkg, oracle, f.cpu().numpy(), name, 0)


@pytest.mark.parametrize("name", ["mt", "sn"])
def test_host_arrays(pkg, oracle, name):
    run_both(pkg, oracle, noise((20, 30, 40), 1, np.float64), name, 0.0, host=True)


@pytest.mark.parametrize("name", ["mt", "sn", "mc"])
def test_gyroid_256_against_oracle(pkg, oracle, name):
    # the tetrahedra / surface-nets workload at its real size: Float32 256^3
    # gyroid on [-2pi, 2pi]^3, iso=0 (MarchingTetrahedra's default eps=1e-3)
    m = {"mt": pkg.MarchingTetrahedra(iso=0), "sn": pkg.NaiveSurfaceNets(iso=0),
         "mc": pkg.MarchingCubes(iso=0)}[name]
    f = pkg.generate_field("gyroid", (256, 256, 256), "float32")
    V, F = pkg.isosurface(f, m)
    kw = dict(iso=0, eps=1e-3) if name == "mt" else dict(iso=0)
    Vo, Fo = oracle.isosurface(f.cpu().numpy(), name, **kw)
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
