"""Shared helpers of the GPU tests (fields, device upload, exact comparison)."""
import numpy as np


def sphere_field(n, radius, dtype):
    r = np.linspace(-1.0, 1.0, n)
    x, y, z = np.meshgrid(r, r, r, indexing="ij")
    return (np.sqrt(x * x + y * y + z * z) - radius).astype(dtype)


def noise(shape, seed, dtype):
    """Smooth-ish random field: white noise blurred along each axis."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape)
    for ax in range(3):
        if f.shape[ax] > 2:
            f = (f + np.roll(f, 1, axis=ax) + np.roll(f, -1, axis=ax)) / 3.0
    return f.astype(dtype)


def dev_field(f):
    """numpy [x,y,z] array -> CUDA tensor [x,y,z] with x fastest in memory."""
    import torch
    t = torch.from_numpy(np.ascontiguousarray(np.asarray(f).transpose(2, 1, 0))).cuda()
    return t.permute(2, 1, 0)


def assert_same_mesh(V, F, Vo, Fo):
    assert V.dtype == Vo.dtype, (V.dtype, Vo.dtype)
    assert V.shape == Vo.shape, (V.shape, Vo.shape)
    assert F.shape == Fo.shape, (F.shape, Fo.shape)
    assert np.array_equal(F, Fo), "faces differ"
    # bit-identical vertices; any NaN matches any NaN
    nan = np.isnan(V)
    assert np.array_equal(nan, np.isnan(Vo)), "NaN positions differ"
    ib = V.view(np.uint32 if V.dtype == np.float32 else np.uint64)
    ob = Vo.view(np.uint32 if Vo.dtype == np.float32 else np.uint64)
    bad = (ib != ob) & ~nan
    if bad.any():
        k = np.argwhere(bad)[0]
        raise AssertionError(f"{int(bad.sum())} vertex coordinates differ; first at {k}: "
                             f"{V[k[0]]} vs {Vo[k[0]]}")
