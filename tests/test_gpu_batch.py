"""GPU parity of the batched form: every volume of a mixed batch (sizes,
element types, methods, iso/eps/axis types, host or device) must equal the
CPU oracle -- and the single-volume path -- bit for bit."""
import numpy as np
import pytest

from helpers import assert_same_mesh, dev_field, noise, sphere_field

pytestmark = pytest.mark.gpu


def oracle_run(oracle, f, m, ax):
    name = {0: "mc", 1: "mt", 2: "sn"}[m.method]
    kw = dict(iso=m.iso)
    if name == "mt":
        kw["eps"] = m.eps
    if ax is not None:
        kw.update(X=ax[0], Y=ax[1], Z=ax[2])
    return oracle.isosurface(f, name, **kw)


def mixed_cases(pkg):
    f32, f64 = np.float32, np.float64
    return [
        (sphere_field(33, 0.5, f32), pkg.MarchingCubes(iso=0), None),
        (sphere_field(20, 0.45, f64), pkg.MarchingTetrahedra(iso=0.0), None),
        (noise((17, 9, 23), 1, f32), pkg.NaiveSurfaceNets(iso=np.float32(0.05)), None),
        (noise((40, 31, 7), 2, f64), pkg.MarchingCubes(iso=0.1),
         ((0, 10), (np.float32(-0.3), np.float32(0.7)), (-2.5, 3.0))),
        (noise((2, 2, 2), 3, f32), pkg.MarchingCubes(iso=0.0), None),
        (noise((1, 5, 5), 4, f32), pkg.MarchingCubes(iso=0.0), ((0.0, 0.0), (-1, 1), (-1, 1))),
        (noise((25, 26, 27), 5, f32), pkg.MarchingTetrahedra(iso=np.float32(0), eps=np.float32(1e-3)),
         ((np.float32(-1), np.float32(1)),) * 3),
        (noise((64, 3, 70), 6, f64), pkg.NaiveSurfaceNets(iso=0), ((0, 1), (0, 1), (0, 1))),
        (sphere_field(47, 0.55, f32), pkg.MarchingCubes(iso=np.float32(0)),
         ((np.float32(-1), np.float32(1)),) * 3),
        (noise((61, 19, 33), 7, f32), pkg.MarchingCubes(iso=0), None),
        (noise((9, 70, 12), 8, f64), pkg.MarchingTetrahedra(iso=0, eps=0), None),
    ]


@pytest.mark.parametrize("where", ["device", "host"])
def test_mixed_batch_matches_oracle(pkg, oracle, where):
    cases = mixed_cases(pkg)
    fields = [dev_field(f) if where == "device" else f for f, _, _ in cases]
    res = pkg.isosurface_batch(fields, [m for _, m, _ in cases], [a for _, _, a in cases])
    assert len(res) == len(cases)
    for (f, m, ax), (V, F) in zip(cases, res):
        if where == "device":
            V, F = V.cpu().numpy(), F.cpu().numpy()
        Vo, Fo = oracle_run(oracle, f, m, ax)
        assert_same_mesh(V, F, Vo, Fo)


def test_batch_equals_single_path_on_device_fields(pkg):
    import torch
    rng = np.random.default_rng(3)
    kinds = ["sphere", "torus", "gyroid", "blobs"]
    fields, methods = [], []
    for v in range(48):
        shape = tuple(int(s) for s in rng.integers(32, 97, size=3))
        dt = "float32" if v % 5 else "float64"
        fields.append(pkg.generate_field(kinds[v % 4], shape, dt, seed=v))
        methods.append([pkg.MarchingCubes(iso=0), pkg.MarchingTetrahedra(iso=0),
                        pkg.NaiveSurfaceNets(iso=0)][v % 3])
    res = pkg.isosurface_batch(fields, methods)
    for f, m, (V, F) in zip(fields, methods, res):
        Vs, Fs = pkg.isosurface(f, m)
        assert V.dtype == Vs.dtype and V.shape == Vs.shape and F.shape == Fs.shape
        assert torch.equal(F, Fs)
        assert torch.equal(V.view(torch.int32 if V.dtype == torch.float32 else torch.int64),
                           Vs.view(torch.int32 if Vs.dtype == torch.float32 else torch.int64))


def test_batches_in_a_row_and_state(pkg, oracle):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    small = [noise((12, 13, 14), s, np.float32) for s in range(3)]
    big = [noise((70, 65, 60), 10 + s, np.float64) for s in range(2)]
    for fields in (small, big, small):
        res = pkg.isosurface_batch([dev_field(f) for f in fields], pkg.MarchingCubes(iso=0))
        for f, (V, F) in zip(fields, res):
            Vo, Fo = oracle.isosurface(f, "mc", iso=0)
            assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
    # an extract that does not follow its count is refused
    import ctypes
    L = capi.lib()
    ctx = capi.context(0)
    f0, keep, _ = pkg._field(small[0])
    p0 = pkg._params(pkg.MarchingCubes(iso=0), (-1, 1), (-1, 1), (-1, 1))
    Fa = (capi.MshField * 1)(f0)
    Pa = (capi.MshParams * 1)(p0)
    nv, nf = (ctypes.c_int64 * 1)(), (ctypes.c_int64 * 1)()
    with ctx.lock:
        assert L.msh_batch_count(ctx.handle, 1, Fa, Pa, nv, nf) == 0
        a, b = ctypes.c_int64(), ctypes.c_int64()
        assert L.msh_count(ctx.handle, ctypes.byref(f0), ctypes.byref(p0), ctypes.byref(a),
                           ctypes.byref(b)) == 0
        assert (a.value, b.value) == (nv[0], nf[0])
        vp, fp = (ctypes.c_void_p * 1)(), (ctypes.c_void_p * 1)()
        assert L.msh_batch_extract(ctx.handle, 1, Fa, Pa, vp, fp, capi.MSH_HOST) == capi.MSH_ERR_STATE


def test_empty_batch(pkg):
    assert pkg.isosurface_batch([]) == []


def _graph_launches(pkg):
    import ctypes
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    ctx = capi.context(0)
    a, b = ctypes.c_int64(), ctypes.c_int64()
    with ctx.lock:
        ctx.check(capi.lib().msh_batch_graph_launches(ctx.handle, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value


def test_repeated_batch_replays_graphs(pkg):
    # A batch that repeats runs from captured hipGraphs (second call
    # captures, later calls replay).  Every replay, with eager library work
    # and synchronisations in between, must equal the single-volume results;
    # a changed batch must not reuse the graphs.
    import torch
    rng = np.random.default_rng(11)
    kinds = ["sphere", "torus", "gyroid", "blobs", "value_noise"]
    fields, methods = [], []
    for v in range(14):
        shape = tuple(int(s) for s in rng.integers(20, 90, size=3))
        fields.append(pkg.generate_field(kinds[v % 5], shape, "float64" if v % 4 == 0 else "float32",
                                         seed=v))
        methods.append([pkg.MarchingCubes(iso=0), pkg.MarchingTetrahedra(iso=0),
                        pkg.NaiveSurfaceNets(iso=0)][v % 3])
    want = [pkg.isosurface(f, m) for f, m in zip(fields, methods)]
    other = pkg.generate_field("gyroid", (50, 40, 30), "float32")
    other_want = pkg.isosurface(other, pkg.MarchingTetrahedra(iso=0))
    torch.cuda.synchronize()
    c0, e0 = _graph_launches(pkg)
    s = torch.cuda.Stream()  # graphs are captured on a stream other than the null stream
    for rep in range(6):
        with torch.cuda.stream(s):
            res = pkg.isosurface_batch(fields, methods)
        s.synchronize()
        for (V, F), (Vs, Fs) in zip(res, want):
            assert torch.equal(F, Fs), rep
            ib = torch.int32 if V.dtype == torch.float32 else torch.int64
            assert torch.equal(V.view(ib), Vs.view(ib)), rep
        # eager work between replays: a single extraction and torch ops
        V2, F2 = pkg.isosurface(other, pkg.MarchingTetrahedra(iso=0))
        assert torch.equal(F2, other_want[1])
        del res
        torch.cuda.synchronize()
    c1, e1 = _graph_launches(pkg)
    assert c1 - c0 >= 4 and e1 - e0 >= 4, (c1 - c0, e1 - e0)
    # a different batch (one field changed) runs eagerly and correctly
    fields2 = list(fields)
    fields2[3] = pkg.generate_field("sphere", (33, 34, 35), "float32")
    want2 = pkg.isosurface(fields2[3], methods[3])
    with torch.cuda.stream(s):
        res2 = pkg.isosurface_batch(fields2, methods)
    s.synchronize()
    assert torch.equal(res2[3][1], want2[1])
    assert torch.equal(res2[0][1], want[0][1])
    c2, e2 = _graph_launches(pkg)
    assert (c2, e2) == (c1, e1)


def test_batch_workload_sizes_equal_single_calls(pkg):
    # volumes drawn as in the batched workload (axes iid from 32..256,
    # sphere / torus / gyroid / blobs, methods mixed, Float32), enough of
    # them to span the size classes; each equals its single call
    import torch
    rng = np.random.default_rng(7)
    kinds = ["sphere", "torus", "gyroid", "blobs"]
    meths = [pkg.MarchingCubes(iso=0), pkg.MarchingTetrahedra(iso=0), pkg.NaiveSurfaceNets(iso=0)]
    fields, methods = [], []
    for v in range(40):
        shape = tuple(int(x) for x in rng.integers(32, 257, size=3))
        fields.append(pkg.generate_field(kinds[int(rng.integers(4))], shape, "float32", seed=v))
        methods.append(meths[int(rng.integers(3))])
    res = pkg.isosurface_batch(fields, methods)
    for f, m, (V, F) in zip(fields, methods, res):
        Vs, Fs = pkg.isosurface(f, m)
        assert V.shape == Vs.shape and torch.equal(F, Fs)
        ib = torch.int32 if V.dtype == torch.float32 else torch.int64
        assert torch.equal(V.view(ib), Vs.view(ib))
