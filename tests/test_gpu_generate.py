"""The device field generator (msh_generate_field) against the numpy
statement of each field's definition (tests/field_defs.py).

Sphere, torus and value noise use only correctly rounded operations
(+, -, *, /, sqrt, floor, integer hashing), so the device values must be
bit-identical to the definitions; the gyroid and the blobs go through
cos/sin/exp, where the device library and numpy may differ by an ulp or
two (tolerance 8e-15 absolute on values of magnitude <= 3)."""
import numpy as np
import pytest

import field_defs

pytestmark = pytest.mark.gpu

EXACT = ("sphere", "torus", "value_noise")
SHAPES = [(37, 29, 23), (64, 64, 64), (2, 3, 130), (1, 5, 6)]


def _check(kind, got, want, dtype):
    if dtype == "float64":
        if kind in EXACT:
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), kind
        else:
            assert np.max(np.abs(got - want)) <= 8e-15, kind
    else:
        w32 = want.astype(np.float32)
        if kind in EXACT:
            assert np.array_equal(got.view(np.uint32), w32.view(np.uint32)), kind
        else:
            # the Float64 value differs by ulps of Float64, so its Float32
            # rounding differs at most by one Float32 ulp
            assert np.all(np.abs(got.astype(np.float64) - w32) <=
                          np.spacing(np.abs(w32)).astype(np.float64)), kind


@pytest.mark.parametrize("dtype", ["float64", "float32"])
@pytest.mark.parametrize("kind", ["sphere", "torus", "gyroid", "value_noise", "blobs"])
def test_generator_matches_definition(pkg, kind, dtype):
    for shape in SHAPES:
        for seed in (0, 7):
            f = pkg.generate_field(kind, shape, dtype, seed=seed)
            assert tuple(f.shape) == shape and f.stride(0) == 1
            got = f.cpu().numpy()
            _check(kind, got, field_defs.field(kind, shape, seed=seed), dtype)


def test_generator_parameters(pkg):
    # explicit p0/p1 reach the kernel
    f = pkg.generate_field("torus", (33, 31, 29), "float64", p0=0.5, p1=0.3)
    _check("torus", f.cpu().numpy(), field_defs.field("torus", (33, 31, 29), p0=0.5, p1=0.3),
           "float64")
    g = pkg.generate_field("gyroid", (20, 21, 22), "float64", p0=3.0)
    _check("gyroid", g.cpu().numpy(), field_defs.field("gyroid", (20, 21, 22), p0=3.0), "float64")


def test_value_noise_headline_planes(pkg):
    # the 1024^3 bench field: its first, a middle and its last x-plane
    n = 1024
    f = pkg.generate_field("value_noise", (n, n, n), "float32", seed=1)
    for x in (0, 517, n - 1):
        got = f[x].cpu().numpy()
        want = field_defs.value_noise_at((n, n, n), 1, [[x], np.arange(n), np.arange(n)])[0]
        assert np.array_equal(got.view(np.uint32), want.astype(np.float32).view(np.uint32)), x
    del f
