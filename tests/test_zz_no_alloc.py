"""msh_isosurface_noalloc, the entry point that allocates nothing, against
msh_count + msh_extract (the entry points it stands for), called through
the C ABI: bit-identical meshes on consecutive calls (the second smaller,
so stale scratch contents from the first would show), short output
capacities, volumes without cells, and the errors it raises."""
import ctypes

import numpy as np
import pytest

from helpers import dev_field

pytestmark = pytest.mark.gpu


def _lib(pkg):
    from importlib import import_module
    capi = import_module(pkg.__name__ + "._capi")
    return capi, capi.lib()


def noalloc(pkg, field, method, vcap=None, fcap=None, axes=((-1, 1),) * 3):
    """Call msh_isosurface_noalloc; returns (V, F, sizes) with V/F of the
    given capacities (default: the exact sizes from msh_count)."""
    import torch
    capi, L = _lib(pkg)
    f, view, _ = pkg._field(field)
    p = pkg._params(method, *axes)
    ctx = capi.context(0)
    if vcap is None or fcap is None:
        nv, nf = pkg.count(field, method, *axes)
        vcap = nv if vcap is None else vcap
        fcap = nf if fcap is None else fcap
    ok = ctypes.c_int32()
    ctx.check(L.msh_output_kind(ctypes.byref(f), ctypes.byref(p), ctypes.byref(ok)))
    dt = torch.float32 if ok.value == capi.MSH_FLOAT32 else torch.float64
    V = torch.full((vcap, 3), float("nan"), dtype=dt, device="cuda")
    F = torch.full((fcap, 3), -7, dtype=torch.int64, device="cuda")
    sizes = torch.full((2,), -1, dtype=torch.int64, device="cuda")
    with ctx.lock:
        ctx.check(L.msh_set_stream(ctx.handle,
                                   ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        ctx.check(L.msh_isosurface_noalloc(ctx.handle, ctypes.byref(f), ctypes.byref(p),
                                           ctypes.c_void_p(V.data_ptr() if vcap else None), vcap,
                                           ctypes.c_void_p(F.data_ptr() if fcap else None), fcap,
                                           ctypes.c_void_p(sizes.data_ptr())))
    torch.cuda.synchronize()
    return V, F, (int(sizes[0]), int(sizes[1]))


def same(got, want):
    import torch
    V, F = got
    Vw, Fw = want
    assert V.dtype == Vw.dtype and V.shape == Vw.shape and F.shape == Fw.shape
    assert torch.equal(F, Fw)
    ib = torch.int32 if V.dtype == torch.float32 else torch.int64
    assert torch.equal(V.view(ib), Vw.view(ib))


@pytest.mark.parametrize("method", ["mc", "mt", "sn"])
def test_noalloc_equals_count_extract_twice(pkg, method):
    m = {"mc": pkg.MarchingCubes(iso=0), "mt": pkg.MarchingTetrahedra(iso=0),
         "sn": pkg.NaiveSurfaceNets(iso=0)}[method]
    big = pkg.generate_field("value_noise", (96, 90, 85), "float32", seed=3)
    small = pkg.generate_field("value_noise", (40, 37, 33), "float64", seed=4)
    for fld in (big, small):
        print(f"noalloc {method} {tuple(fld.shape)} {fld.dtype}", flush=True)
        V, F, sz = noalloc(pkg, fld, m)
        want = pkg.isosurface(fld, m)
        assert sz == (len(want[0]), len(want[1]))
        same((V, F), want)


def test_noalloc_headline_plane_count(pkg):
    # the largest volumes the scratch takes: 1024 x 1024 planes
    fld = pkg.generate_field("value_noise", (1024, 1024, 40), "float32", seed=1)
    m = pkg.MarchingCubes(iso=0)
    V, F, sz = noalloc(pkg, fld, m)
    same((V, F), pkg.isosurface(fld, m))
    # taller, with tetrahedra: the other entry points cut these z rows into
    # four segments (four times the workspace, past the scratch); this one
    # keeps them whole, and the meshes agree
    fld = pkg.generate_field("sphere", (1024, 1024, 300), "float32")
    mt = pkg.MarchingTetrahedra(iso=0)
    V, F, sz = noalloc(pkg, fld, mt)
    same((V, F), pkg.isosurface(fld, mt))
    del fld
    # and a small one right after it
    fld = pkg.generate_field("sphere", (20, 21, 22), "float32")
    V, F, sz = noalloc(pkg, fld, m)
    same((V, F), pkg.isosurface(fld, m))


def test_noalloc_short_capacity(pkg):
    import torch
    fld = pkg.generate_field("gyroid", (48, 50, 52), "float32")
    m = pkg.MarchingCubes(iso=0)
    Vw, Fw = pkg.isosurface(fld, m)
    nv, nf = len(Vw), len(Fw)
    V, F, sz = noalloc(pkg, fld, m, vcap=nv // 3, fcap=nf // 2)
    assert sz == (nv, nf)  # the exact sizes, so the caller can retry
    same((V, F), (Vw[:nv // 3], Fw[:nf // 2]))
    V, F, sz = noalloc(pkg, fld, m, vcap=0, fcap=0)
    assert sz == (nv, nf) and V.numel() == 0 and F.numel() == 0


def test_noalloc_no_cells(pkg):
    import torch
    # one x-sample: no cells (a length-1 axis needs first == last, as
    # Julia's LinRange does)
    fld = dev_field(np.zeros((1, 5, 6), np.float32))
    V, F, sz = noalloc(pkg, fld, pkg.MarchingCubes(iso=0), vcap=4, fcap=4,
                       axes=((0, 0), (-1, 1), (-1, 1)))
    assert sz == (0, 0)
    assert bool(torch.isnan(V).all()) and bool((F == -7).all())


def test_noalloc_errors(pkg):
    import torch
    capi, L = _lib(pkg)
    ctx = capi.context(0)
    m = pkg.MarchingCubes(iso=0)
    p = pkg._params(m, (-1, 1), (-1, 1), (-1, 1))
    sizes = torch.zeros(2, dtype=torch.int64, device="cuda")
    # workspace beyond the static scratch: refused before any launch
    t = torch.zeros((4, 2048, 2048), dtype=torch.float32, device="cuda").permute(2, 1, 0)
    f, _, _ = pkg._field(t)
    rc = L.msh_isosurface_noalloc(ctx.handle, ctypes.byref(f), ctypes.byref(p), None, 0, None, 0,
                                  ctypes.c_void_p(sizes.data_ptr()))
    assert rc == capi.MSH_ERR_ARGUMENT
    assert b"static scratch" in L.msh_last_error(ctx.handle)
    # host fields are refused
    h = np.asfortranarray(np.zeros((4, 4, 4), np.float32))
    fh = capi.MshField(capi.MSH_FLOAT32, capi.MSH_HOST, 4, 4, 4,
                       ctypes.c_void_p(h.ctypes.data))
    rc = L.msh_isosurface_noalloc(ctx.handle, ctypes.byref(fh), ctypes.byref(p), None, 0, None, 0,
                                  ctypes.c_void_p(sizes.data_ptr()))
    assert rc == capi.MSH_ERR_ARGUMENT and b"device field" in L.msh_last_error(ctx.handle)
    torch.cuda.synchronize()
