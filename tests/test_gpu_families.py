"""Every emit kernel family (lanes along x, lanes along z, generic) gives the
oracle's mesh bit for bit, through the single and the batched entry points,
and the autotuner picks one of them without changing results."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, noise, sphere_field

pytestmark = pytest.mark.gpu

CASES = [("mc", np.float32), ("mc", np.float64), ("mt", np.float32), ("mt", np.float64),
         ("sn", np.float32), ("sn", np.float64)]


def method_of(pkg, name):
    return {"mc": pkg.MarchingCubes(iso=0), "mt": pkg.MarchingTetrahedra(iso=0),
            "sn": pkg.NaiveSurfaceNets(iso=0)}[name]


@pytest.fixture
def reset_families(pkg):
    yield
    for name, dt in CASES:
        pkg.set_kernel_family(method_of(pkg, name), dt, -1)


@pytest.mark.parametrize("family", [3, 2, 1, 0])
@pytest.mark.parametrize("name,dtype", CASES)
def test_family_matches_oracle(pkg, oracle, reset_families, family, name, dtype):
    m = method_of(pkg, name)
    pkg.set_kernel_family(m, dtype, family)
    fields = [noise((37, 29, 41), 3, dtype), sphere_field(30, 0.55, dtype)]
    for f in fields:
        V, F = pkg.isosurface(dev_field(f), m)
        Vo, Fo = oracle.isosurface(f, name, iso=0)
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
    res = pkg.isosurface_batch([dev_field(f) for f in fields], m)
    for f, (V, F) in zip(fields, res):
        Vo, Fo = oracle.isosurface(f, name, iso=0)
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)


def sparse_field(dtype):
    # a small sphere off centre in a larger ragged box: most emit tiles of
    # every family hold no active cell
    x, y, z = np.meshgrid(np.linspace(-1, 1, 150), np.linspace(-1, 1, 70),
                          np.linspace(-1, 1, 90), indexing="ij")
    return (np.sqrt((x - 0.4) ** 2 + (y + 0.3) ** 2 + z * z) - 0.22).astype(dtype)


@pytest.mark.parametrize("family", [3, 2, 1, 0])
@pytest.mark.parametrize("name,dtype", CASES)
def test_family_skips_empty_tiles(pkg, oracle, reset_families, family, name, dtype):
    # tiles whose rows the count found empty are skipped whole; the meshes
    # (alone, and in a batch with a volume nothing crosses) do not change
    m = method_of(pkg, name)
    pkg.set_kernel_family(m, dtype, family)
    f = sparse_field(dtype)
    empty = np.ones((40, 33, 20), dtype)
    Vo, Fo = oracle.isosurface(f, name, iso=0)
    assert len(Fo) > 1000
    V, F = pkg.isosurface(dev_field(f), m)
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
    res = pkg.isosurface_batch([dev_field(f), dev_field(empty), dev_field(f[::-1].copy())], m)
    assert len(res[1][1]) == 0 and len(res[1][0]) == 0
    assert_same_mesh(res[0][0].cpu().numpy(), res[0][1].cpu().numpy(), Vo, Fo)
    Vr, Fr = oracle.isosurface(f[::-1].copy(), name, iso=0)
    assert_same_mesh(res[2][0].cpu().numpy(), res[2][1].cpu().numpy(), Vr, Fr)


def tall_field(nz, dtype):
    # a thin column with a few separated blobs along z: in a batch its z
    # rows are one segment of nz - 1 cells, so the z-block masks use blocks
    # of 32 (segments up to 2048 cells) or 64 cells (up to 4096)
    x, y, z = np.meshgrid(np.linspace(-1, 1, 24), np.linspace(-1, 1, 20),
                          np.linspace(-1, 1, nz), indexing="ij")
    f = np.full(x.shape, 1.0)
    for zc in (-0.9, -0.2, 0.05, 0.7):
        f = np.minimum(f, np.sqrt(x * x + y * y + ((z - zc) * 20) ** 2) - 0.5)
    return f.astype(dtype)


@pytest.mark.parametrize("family", [3, 2])
@pytest.mark.parametrize("name,dtype", CASES)
def test_family_long_z_segments(pkg, oracle, reset_families, family, name, dtype):
    m = method_of(pkg, name)
    pkg.set_kernel_family(m, dtype, family)
    fields = [tall_field(1500, dtype), tall_field(3000, dtype)]
    res = pkg.isosurface_batch([dev_field(f) for f in fields], m)
    for f, (V, F) in zip(fields, res):
        Vo, Fo = oracle.isosurface(f, name, iso=0)
        assert len(Fo) > 100
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
        V1, F1 = pkg.isosurface(dev_field(f), m)  # single volume: shorter z segments
        assert_same_mesh(V1.cpu().numpy(), F1.cpu().numpy(), Vo, Fo)


@pytest.mark.parametrize("name,dtype", CASES)
def test_autotune_keeps_results(pkg, oracle, reset_families, name, dtype):
    m = method_of(pkg, name)
    f = noise((70, 66, 60), 5, dtype)
    fam = pkg.autotune(dev_field(f), m)
    assert fam in (0, 1, 2, 3)
    V, F = pkg.isosurface(dev_field(f), m)
    Vo, Fo = oracle.isosurface(f, name, iso=0)
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)


def test_autotune_batch_keeps_results(pkg, oracle, reset_families):
    # the batch autotuner times each (method, element type, size class)
    # group as a batched launch with every family; results do not change
    rng = np.random.default_rng(21)
    cases = []
    for v in range(12):
        shape = tuple(int(s) for s in rng.integers(12, 48, size=3))
        dt = np.float64 if v % 4 == 0 else np.float32
        cases.append((noise(shape, 40 + v, dt), ["mc", "mt", "sn"][v % 3]))
    fields = [dev_field(f) for f, _ in cases]
    methods = [method_of(pkg, name) for _, name in cases]
    chosen = pkg.autotune_batch(fields, methods)
    assert len(chosen) == len(cases) and all(c in (-1, 0, 1, 2, 3) for c in chosen)
    res = pkg.isosurface_batch(fields, methods)
    for (f, name), (V, F) in zip(cases, res):
        Vo, Fo = oracle.isosurface(f, name, iso=0)
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)


def waves(shape, seed, dtype):
    # smooth: long runs of cells without a crossing, so whole z blocks of a
    # row group stay unmarked
    rng = np.random.default_rng(seed)
    x, y, z = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij", sparse=True)
    k = rng.uniform(1, 4, 3) * np.pi
    return (np.cos(k[0] * x) * np.sin(k[1] * y) + np.cos(k[1] * y) * np.sin(k[2] * z) +
            np.cos(k[2] * z) * np.sin(k[0] * x)).astype(dtype)


@pytest.mark.parametrize("name,dtype", CASES)
def test_count_split_over_z(pkg, oracle, name, dtype):
    # a volume too small to fill the device on its own has 2 or 4 count waves
    # share each x span, each over its own whole z blocks: 4 waves over 64-
    # and 39-cell segments (some waves without blocks) in the first three,
    # 2 waves with a wave beyond the x end in the last tile in the fourth
    # (26 M samples: Float32 only); alone and (no sharing) in a batch
    m = method_of(pkg, name)
    fields = [waves((300, 120, 200), 21, dtype), waves((151, 260, 104), 22, dtype),
              waves((40, 40, 1500), 23, dtype)]
    if dtype == np.float32:
        fields.append(waves((151, 100, 1729), 24, dtype))
    outs = [pkg.isosurface(dev_field(f), m) for f in fields]
    res = pkg.isosurface_batch([dev_field(f) for f in fields], m)
    for f, (V, F), (Vb, Fb) in zip(fields, outs, res):
        Vo, Fo = oracle.isosurface(f, name, iso=0)
        assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
        assert_same_mesh(Vb.cpu().numpy(), Fb.cpu().numpy(), Vo, Fo)
