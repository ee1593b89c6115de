"""The benchmark workloads at their real sizes.

Marching Cubes on the Float64 1024^3 value-noise field (32^3 lattice of
seeded Uniform(-1,1) values trilinearly upsampled, plus 0.25 x a 128^3
octave), iso=0, X=Y=Z=-1:1: 2^30 samples, an 8 GiB field and about 272 M
vertices, so byte offsets run far past 2^32.  The oracle cannot mesh the
whole field in test time, but the reference's loop runs x outermost
(src/marching_cubes.jl:40) and a cube's faces name only its own vertices
(src/marching_cubes.jl:69-90), so the mesh is the concatenation of its
x-slabs' meshes with shifted faces.  The tests compare x-slabs at the start,
in the middle and at the end of the GPU mesh with the oracle run on just
those cells (with the coordinates of the whole 1024-sample axes), and check
properties of the whole mesh that any slip in offsets would break.

A Float32 field and the same samples widened to Float64 must give the same
faces, with vertices within 1.2e-7 of each other (the Float32 output is the
Float64 lerp rounded once to Float32: half an ulp of a coordinate in
[-1, 1] is at most 6e-8).
"""
import numpy as np
import pytest

from helpers import assert_same_mesh

pytestmark = pytest.mark.gpu

N = 1024
SLAB = 8  # x-cells per compared slab


@pytest.fixture(scope="module")
def mesh64(pkg):
    import torch
    print(f"headline: Float64 {N}^3 value noise, MarchingCubes(iso=0)", flush=True)
    f = pkg.generate_field("value_noise", (N, N, N), "float64", seed=1)
    V, F = pkg.isosurface(f, pkg.MarchingCubes(iso=0))
    torch.cuda.synchronize()
    yield f, V, F
    del f, V, F
    torch.cuda.empty_cache()


def slab_cells(oracle, f, xlo, xhi):
    """Oracle mesh of cells xlo..xhi (1-based) of the device field f."""
    slab = f[xlo - 1:xhi + 1].cpu().numpy()  # x-planes xlo..xhi+1
    return oracle.isosurface_xslab(slab, N, xlo - 1, xlo, xhi, iso=0)


def test_headline_shape(mesh64):
    import torch
    f, V, F = mesh64
    assert V.dtype == torch.float64 and F.dtype == torch.int64
    # dense: vertex byte offsets run past 2^32, face byte offsets past 2^31
    assert V.numel() * 8 > 2 ** 32 and F.numel() * 8 > 2 ** 31
    assert len(F) > 10 ** 8


def test_headline_prefix(oracle, mesh64):
    f, V, F = mesh64
    Vo, Fo = slab_cells(oracle, f, 1, SLAB)
    assert len(Fo) > 1000
    assert_same_mesh(V[:len(Vo)].cpu().numpy(), F[:len(Fo)].cpu().numpy(), Vo, Fo)


def test_headline_suffix(oracle, mesh64):
    f, V, F = mesh64
    Vo, Fo = slab_cells(oracle, f, N - SLAB, N - 1)
    assert len(Fo) > 1000
    shift = len(V) - len(Vo)
    assert_same_mesh(V[shift:].cpu().numpy(), (F[len(F) - len(Fo):] - shift).cpu().numpy(),
                     Vo, Fo)


def test_headline_middle(pkg, oracle, mesh64):
    # the vertices and faces before cell x = 513 are what the field's first
    # 513 x-planes give (the counts do not depend on the coordinates)
    f, V, F = mesh64
    x0 = 513
    nv0, nf0 = pkg.count(f[:x0], pkg.MarchingCubes(iso=0))
    Vo, Fo = slab_cells(oracle, f, x0, x0 + SLAB - 1)
    assert_same_mesh(V[nv0:nv0 + len(Vo)].cpu().numpy(),
                     (F[nf0:nf0 + len(Fo)] - nv0).cpu().numpy(), Vo, Fo)


def test_headline_whole_mesh_properties(mesh64):
    import torch
    f, V, F = mesh64
    nv = len(V)
    lo = F.min(dim=1).values
    hi = F.max(dim=1).values
    assert int(lo[0]) == 1 and int(hi.max()) == nv
    # faces come cube by cube in emission order and name only their own
    # cube's vertices (at most 12 per cube)
    assert bool((lo[1:] >= lo[:-1]).all())
    assert bool(((hi - lo) <= 11).all())
    del lo, hi
    # every vertex belongs to some face
    seen = torch.zeros(nv + 1, dtype=torch.bool, device=V.device)
    seen[F.reshape(-1)] = True
    assert bool(seen[1:].all())
    del seen
    # vertices lie in the [-1, 1]^3 box of the axes
    assert float(V.abs().max()) <= 1.0
    assert not bool(torch.isnan(V).any())


def test_headline_float32_matches_float64_of_same_samples(pkg):
    import torch
    f32 = pkg.generate_field("value_noise", (N, N, N), "float32", seed=1)
    V32, F32 = pkg.isosurface(f32, pkg.MarchingCubes(iso=0))
    f64 = f32.permute(2, 1, 0).double().permute(2, 1, 0)  # exact, x fastest
    del f32
    V64, F64 = pkg.isosurface(f64, pkg.MarchingCubes(iso=0))
    del f64
    assert V32.dtype == torch.float32 and V64.dtype == torch.float64
    assert V32.shape == V64.shape
    assert torch.equal(F32, F64)
    err = float((V32.double() - V64).abs().max())
    assert err <= 1.2e-7, err
    del V32, F32, V64, F64
    torch.cuda.empty_cache()


def test_sparse_sphere_1024_float32(pkg, oracle):
    # the sparse 1024^3 point of the benchmark: Float32 sphere of radius 0.5
    # on [-1, 1]^3; the middle x-slab (through the sphere's widest part)
    # against the oracle, its offsets from counting the planes before it
    import torch
    print(f"sparse sphere: Float32 {N}^3, MarchingCubes(iso=0)", flush=True)
    f = pkg.generate_field("sphere", (N, N, N), "float32", p0=0.5)
    V, F = pkg.isosurface(f, pkg.MarchingCubes(iso=0))
    assert V.dtype == torch.float32 and 10 ** 6 < len(F) < 10 ** 7
    x0 = 509
    nv0, nf0 = pkg.count(f[:x0], pkg.MarchingCubes(iso=0))
    Vo, Fo = slab_cells(oracle, f, x0, x0 + SLAB - 1)
    assert len(Fo) > 10000
    assert_same_mesh(V[nv0:nv0 + len(Vo)].cpu().numpy(),
                     (F[nf0:nf0 + len(Fo)] - nv0).cpu().numpy(), Vo, Fo)
    # every vertex within half a cell of the sphere (the surface crosses
    # each cut edge, and an edge is one cell long)
    r = V.double().norm(dim=1)
    assert float((r - 0.5).abs().max()) <= 1.0 / (N - 1)
    del f, V, F, r
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_value_noise_256_against_oracle(pkg, oracle, dtype):
    name = "float64" if dtype == np.float64 else "float32"
    f = pkg.generate_field("value_noise", (256, 256, 256), name, seed=5)
    V, F = pkg.isosurface(f, pkg.MarchingCubes(iso=0))
    Vo, Fo = oracle.isosurface(f.cpu().numpy(), "mc", iso=0)
    assert_same_mesh(V.cpu().numpy(), F.cpu().numpy(), Vo, Fo)
