"""GPU parity of msh_weld (through the C ABI) with the welding definition
in oracle.weld, bit for bit: Marching Cubes and Surface Nets outputs of the
GPU path, heavy duplication (long hash probe chains), bitwise -0.0/NaN
keys, empty meshes, host arrays, bad indices; and at 256^3 the defining
properties."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, noise, sphere_field

pytestmark = pytest.mark.gpu


def weld_both(pkg, oracle, V, F):
    import torch
    Vt = V if isinstance(V, torch.Tensor) else torch.from_numpy(V).cuda()
    Ft = F if isinstance(F, torch.Tensor) else torch.from_numpy(F).cuda()
    Vw, Fw = pkg.weld(Vt, Ft)
    torch.cuda.synchronize()
    Vo, Fo = oracle.weld(Vt.cpu().numpy(), Ft.cpu().numpy())
    assert_same_mesh(Vw.cpu().numpy(), Fw.cpu().numpy(), Vo, Fo)
    return Vw, Fw


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("method", ["mc", "sn", "mt"])
def test_weld_gpu_meshes(pkg, oracle, dtype, method):
    m = {"mc": pkg.MarchingCubes(iso=0.0), "sn": pkg.NaiveSurfaceNets(),
         "mt": pkg.MarchingTetrahedra(iso=0.0)}[method]
    for f in (sphere_field(64, 0.55, dtype), noise((48, 40, 56), 11, dtype)):
        V, F = pkg.isosurface(dev_field(f), m)
        Vw, Fw = weld_both(pkg, oracle, V, F)
        if method == "mc":
            assert Vw.shape[0] < V.shape[0] // 2


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_weld_heavy_duplication(pkg, oracle, dtype):
    rng = np.random.default_rng(3)
    pts = rng.standard_normal((997, 3)).astype(dtype)
    V = pts[rng.integers(0, 997, 200_000)]
    V[::13] = [-0.0, 0.0, 0.5]
    V[::17] = [0.0, 0.0, 0.5]
    V[1::101] = [np.nan, 0.0, 0.5]
    F = rng.integers(1, V.shape[0] + 1, (150_000, 3))
    weld_both(pkg, oracle, V, F)
    # one key repeated everywhere: a single hash slot takes every atomic
    weld_both(pkg, oracle, np.ones((100_000, 3), dtype), F[:1000] % 100_000 + 1)


def test_weld_empty_and_small(pkg, oracle):
    import torch
    weld_both(pkg, oracle, np.zeros((0, 3)), np.zeros((0, 3), np.int64))
    weld_both(pkg, oracle, np.ones((4, 3), np.float32), np.zeros((0, 3), np.int64))
    weld_both(pkg, oracle, np.arange(9.0).reshape(3, 3), np.array([[1, 2, 3], [3, 2, 1]]))
    Vw, Fw = pkg.weld(np.zeros((3, 3)), np.array([[1, 2, 3]]))  # host arrays
    assert isinstance(Vw, np.ndarray) and Vw.shape == (1, 3) and Fw.tolist() == [[1, 1, 1]]
    with pytest.raises(ValueError):
        pkg.weld(torch.zeros((3, 3), device="cuda"),
                 torch.tensor([[1, 2, 4]], device="cuda"))
    with pytest.raises(TypeError):
        pkg.weld(torch.zeros((3, 3), dtype=torch.float16, device="cuda"),
                 torch.tensor([[1, 2, 3]], device="cuda"))
    # the C ABI refuses an output that overlaps the input
    import ctypes
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    V = torch.zeros((4, 3), dtype=torch.float64, device="cuda")
    F = torch.tensor([[1, 2, 3]], device="cuda")
    ctx = capi.context(0)
    n = ctypes.c_int64()
    rc = capi.lib().msh_weld(ctx.handle, capi.MSH_FLOAT64, 4, ctypes.c_void_p(V.data_ptr()), 1,
                             ctypes.c_void_p(F.data_ptr()),
                             ctypes.c_void_p(V.data_ptr() + 24), ctypes.byref(n))
    assert rc == capi.MSH_ERR_ARGUMENT


@pytest.mark.parametrize("dtype", ["float32", "float64"])
def test_weld_value_noise_256(pkg, oracle, dtype):
    import torch
    f = pkg.generate_field("value_noise", (256, 256, 256), dtype, seed=2)
    V, F = pkg.isosurface(f, pkg.MarchingCubes(iso=0))
    Vw, Fw = weld_both(pkg, oracle, V, F)
    # the same again on the same context (grown workspace reused)
    Vw2, Fw2 = pkg.weld(V, F)
    assert torch.equal(Vw.view(torch.uint8), Vw2.view(torch.uint8)) and torch.equal(Fw, Fw2)
    # faces keep their positions
    assert torch.equal(Vw[Fw - 1].view(torch.uint8), V[F - 1].view(torch.uint8))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_weld_large_table_skewed_x(pkg, oracle, dtype):
    # from 2^23 hash slots up, keys are placed by x region (quantiles of a
    # sample of the x coordinates): skewed x, a plane x = const holding half
    # the points, signed zeros, NaNs and an order without any locality
    # still weld to the definition
    rng = np.random.default_rng(7)
    n = 3_500_000  # 1.25 n rounds up to 2^23 slots
    pts = rng.standard_normal((1_200_000, 3)).astype(dtype)
    pts[:600_000, 0] = 0.25
    pts[600_000:700_000, 0] = -0.0
    pts[700_000:800_000, 0] = 0.0
    pts[800_000:810_000, 0] = np.nan
    pts[810_000:820_000, 0] = -np.inf
    V = pts[rng.integers(0, len(pts), n)]
    F = rng.integers(1, n + 1, (1_000_000, 3))
    weld_both(pkg, oracle, V, F)
    # every x the same: every key spreads over all regions by its hash
    V[:, 0] = 1.0
    weld_both(pkg, oracle, V, F)
    # x sorted (the order of an extraction) and 2^24 slots
    V2 = np.concatenate([pts[rng.integers(0, len(pts), 4_000_000)], V[:3_000_000]])
    V2 = V2[np.argsort(V2[:, 0], kind="stable")]
    weld_both(pkg, oracle, V2, rng.integers(1, len(V2) + 1, (500_000, 3)))


def test_weld_batched_meshes(pkg, oracle):
    # every mesh of a mixed batch welds to what the definition gives
    rng = np.random.default_rng(5)
    fields, methods = [], []
    for v in range(12):
        shape = tuple(int(s) for s in rng.integers(8, 40, 3))
        dt = np.float32 if v % 3 else np.float64
        fields.append(dev_field(noise(shape, 100 + v, dt)))
        methods.append([pkg.MarchingCubes(iso=0.0), pkg.NaiveSurfaceNets(),
                        pkg.MarchingTetrahedra(iso=0.0)][v % 3])
    for V, F in pkg.isosurface_batch(fields, methods):
        weld_both(pkg, oracle, V, F)
