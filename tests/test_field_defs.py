"""CPU checks of the numpy field definitions (tests/field_defs.py) before
they are used to check the device generator."""
import numpy as np

import field_defs as d


def test_splitmix64_published_sequence():
    # SplitMix64 from state 0: the generator adds the golden-ratio gamma
    # and mixes; its first three outputs are the published ones.
    g = 0x9E3779B97F4A7C15
    got = [int(d.splitmix64(k * g % 2 ** 64)) for k in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_u01_range_and_resolution():
    assert d.u01(0) == 0.0
    assert d.u01(2 ** 64 - 1) == 1.0 - 2.0 ** -53
    assert d.u01(1 << 11) == 2.0 ** -53


def test_lin_matches_endpoints_and_step():
    r = d.lin(5, -1.0, 1.0)
    assert r.tolist() == [-1.0, -0.5, 0.0, 0.5, 1.0]
    assert d.lin(1, -1.0, 1.0).tolist() == [-1.0]


def test_sphere_and_torus_closed_forms():
    s = d.field("sphere", (5, 5, 5), p0=0.5)
    assert s[2, 2, 2] == -0.5
    assert s[0, 0, 0] == np.sqrt(3.0) - 0.5
    t = d.field("torus", (5, 5, 5), p0=0.6, p1=0.2)
    # on the ring's plane at distance 1 from the axis: |1 - 0.6| - 0.2
    assert np.isclose(t[4, 2, 2], 0.2)
    assert np.isclose(t[2, 2, 2], 0.4)  # the axis: sqrt(0.6^2) - 0.2


def test_gyroid_values():
    g = d.field("gyroid", (3, 3, 3), p0=np.pi / 2)
    # x, y, z in {-pi/2, 0, pi/2}
    assert np.isclose(g[1, 2, 1], 1.0)      # (0, pi/2, 0): cos 0 sin(pi/2)
    assert np.isclose(g[2, 2, 2], 0.0, atol=1e-15)
    assert np.isclose(g[0, 2, 1], -1.0)     # (-pi/2, pi/2, 0): cos 0 sin(-pi/2)


def test_value_noise_interpolates_the_lattice():
    # samples that fall on lattice points take the lattice value itself,
    # and a sample halfway between two lattice points along one axis is
    # their mean
    L, seed = 4, 99
    shape = (7, 4, 4)  # x spacing: half a lattice cell; y, z: on the lattice
    v = d.value_noise(shape, L, seed)

    def lat(a, b, c):
        li = np.uint64(a + L * (b + L * c))
        with np.errstate(over="ignore"):
            h = d.splitmix64(np.uint64(seed) ^ (li * np.uint64(0xD1B54A32D192ED03)))
        return 2.0 * float(d.u01(h)) - 1.0

    for a in range(4):
        for b in range(4):
            for c in range(4):
                assert v[2 * a, b, c] == lat(a, b, c)
    for a in range(3):
        assert np.isclose(v[2 * a + 1, 1, 2], 0.5 * (lat(a, 1, 2) + lat(a + 1, 1, 2)))
    full = d.field("value_noise", (33, 20, 17), seed=5)
    assert np.all(np.abs(full) <= 1.25)


def test_blobs_peak_at_centres():
    f = d.field("blobs", (41, 41, 41), seed=3)
    assert f.max() <= 0.5
    assert f.min() < -0.3  # a Gaussian's peak is 1 above the 0.5 offset
