"""CPU checks that pin the oracle (oracle/cpu_ref.cpp) before it is trusted.

Each check is chosen so that a plausible slip in the restatement fails it:
the reference's own known-answer tests (test/runtests.jl), Julia Base's
documented LinRange values, closed forms on single cubes, and invariants
of closed surfaces (Euler characteristic, enclosed volume, orientation).
"""
import itertools

import numpy as np
import pytest

# MC corner order (src/marching_cubes.jl:42-49) and MT corner order
# (src/marching_tetrahedra.jl:146-153), as (dx, dy, dz).
MC_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
MT_CORNERS = [(0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1), (1, 0, 1)]


def grid(n, lo=-1.0, hi=1.0):
    r = np.linspace(lo, hi, n)
    return np.meshgrid(r, r, r, indexing="ij")


def sphere(n, radius=0.5, dtype=np.float64):
    x, y, z = grid(n)
    return (np.sqrt(x * x + y * y + z * z) - radius).astype(dtype)


def signed_volume(V, F):
    V = V.astype(np.float64)
    a, b, c = V[F[:, 0] - 1], V[F[:, 1] - 1], V[F[:, 2] - 1]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def euler(F):
    """V - E + F of a shared-vertex triangle mesh."""
    edges = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    ne = len(np.unique(edges, axis=0))
    return len(np.unique(F)) - ne + len(F)


def directed_edges_paired(F):
    """Every directed edge of a closed oriented mesh appears once, and its
    reverse appears once."""
    d = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
    fw = {tuple(e) for e in d}
    return len(fw) == len(d) and all((b, a) in fw for a, b in fw)


# ---- known-answer tests of the reference (test/runtests.jl) ---------------

def test_cubeindex_kat(oracle):
    # test/runtests.jl:15-27
    assert oracle.cubeindex([0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8], 0.35) == 0x07
    assert oracle.cubeindex([0.5, 0.6, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2], 0.75) == 0x07
    assert oracle.cubeindex([0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2], 0.5) == 0xE0


def test_cubeindex_strict_and_nan(oracle):
    # `<` is strict and NaN compares false (src/common.jl:10-20)
    assert oracle.cubeindex([0.5] * 8, 0.5) == 0
    assert oracle.cubeindex([np.nan] * 8, 0.5) == 0
    assert oracle.cubeindex([-1.0] * 8, 0.0) == 0xFF
    for bit in range(8):
        v = [1.0] * 8
        v[bit] = -1.0
        assert oracle.cubeindex(v, 0.0) == 1 << bit


def test_vertex_interp_kat(oracle):
    # test/runtests.jl:29-33
    assert oracle.vertex_interp_int(0, (0, 0, 0), (0, 1, 0), -1, 1) == (0, 0.5, 0)
    assert oracle.vertex_interp_int(-1, (0, 0, 0), (0, 1, 0), -1, 1) == (0, 0, 0)
    assert oracle.vertex_interp_int(1, (0, 0, 0), (0, 1, 0), -1, 1) == (0, 1, 0)


def test_linrange_matches_julia_documentation(oracle):
    # Julia Base's LinRange docstring (base/range.jl, not part of the
    # reference) shows these exact values; they pin lerpi's rounding.
    got = oracle.linrange(-0.1, 0.3, 5)
    assert got.tolist() == [-0.1, -1.3877787807814457e-17, 0.09999999999999999,
                            0.19999999999999998, 0.3]
    assert oracle.linrange(1.5, 5.5, 9).tolist() == [1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 5.0, 5.5]
    assert oracle.linrange(-1, 1, 3).tolist() == [-1.0, 0.0, 1.0]


def test_linrange_docstring_rejects_other_roundings():
    # The docstring values above tell lerpi's (1-t)*a + t*b apart from the
    # other usual formulas: a + t*(b-a) and a + i*step (numpy's linspace)
    # both give 0.0 for the second value, not -1.3877787807814457e-17.
    a, b, n = -0.1, 0.3, 5
    t = np.arange(n) / (n - 1)
    assert ((1 - t) * a + t * b)[1] == -1.3877787807814457e-17
    assert (a + t * (b - a))[1] == 0.0
    assert np.linspace(a, b, n)[1] == 0.0


def test_linrange_endpoints_and_float32(oracle):
    for n in (2, 3, 7, 128, 1000):
        r = oracle.linrange(-1, 1, n)
        assert r[0] == -1.0 and r[-1] == 1.0
        assert np.all(np.diff(r) > 0)
    r = oracle.linrange(np.float32(-1.3), np.float32(2.1), 11)
    # Float32 ends: values are Float32 (stored widened)
    assert np.array_equal(r, r.astype(np.float32).astype(np.float64))
    assert r[0] == np.float32(-1.3) and r[-1] == np.float32(2.1)


def test_linrange_length_one(oracle):
    assert oracle.linrange(2.0, 2.0, 1).tolist() == [2.0]
    with pytest.raises(ValueError, match="endpoints differ"):
        oracle.linrange(-1, 1, 1)


@pytest.mark.parametrize("method,iso", [("mc", 0.5), ("mt", 0.5), ("sn", 0.5)])
def test_respect_origin(oracle, method, iso):
    # test/runtests.jl:127-149 on norm_sdf = [norm((x,y,z)) for x,y,z in -1:0.01:1]
    x, y, z = grid(201)
    norm_sdf = np.sqrt(x * x + y * y + z * z)
    V, F = oracle.isosurface(norm_sdf, method, iso=iso)
    assert np.allclose(V.mean(axis=0), 0, atol=0.015)
    if method != "sn":  # SN vertices sit inside cells, not on edges
        for i in range(3):
            assert abs(V[:, i].max() - 0.5) <= 0.001
            assert abs(V[:, i].min() + 0.5) <= 0.001


@pytest.mark.parametrize("method", ["mc", "mt"])
def test_mean_distance_equals_iso(oracle, method):
    # test/runtests.jl:105-120 (mean norm of vertices ~ isovalue), on a
    # coarser grid than the reference's -2:0.01:2 to keep the CPU suite short
    r = np.arange(-200, 201) * 0.01
    x, y, z = np.meshgrid(r[::2], r[::2], r[::2], indexing="ij")
    sdf = np.sqrt(x * x + y * y + z * z)
    V, _ = oracle.isosurface(sdf, method, iso=0.85, X=(-2.0, 2.0), Y=(-2.0, 2.0), Z=(-2.0, 2.0))
    assert abs(np.linalg.norm(V, axis=1).mean() - 0.85) < 1e-2


def test_mixed_types_mt(oracle):
    # test/runtests.jl:81-97: Float64 and Float32 samples give equal counts
    rng = np.random.default_rng(9)
    for _ in range(5):
        s = rng.standard_normal((10, 10, 10))
        V1, F1 = oracle.isosurface(s, "mt", X=(0, 1), Y=(0, 1), Z=(0, 1))
        V2, F2 = oracle.isosurface(s.astype(np.float32), "mt", X=(0, 1), Y=(0, 1), Z=(0, 1))
        assert len(V1) == len(V2)
        assert np.array_equal(F1, F2)


def test_output_types(oracle):
    # float(FT) with FT = promote_type(axes, T, typeof(iso)[, typeof(eps)])
    # (src/marching_cubes.jl:31, src/marching_tetrahedra.jl:131;
    # test/runtests.jl:35-75)
    s = sphere(8, dtype=np.float32)
    assert oracle.isosurface(s, "mc", iso=0)[0].dtype == np.float32
    assert oracle.isosurface(s, "mc", iso=0.0)[0].dtype == np.float64
    assert oracle.isosurface(s, "mc", iso=np.float32(0))[0].dtype == np.float32
    assert oracle.isosurface(s, "mc", iso=0, X=(-1.0, 1.0))[0].dtype == np.float64
    assert oracle.isosurface(s, "mt", iso=0, eps=1e-3)[0].dtype == np.float64
    assert oracle.isosurface(s, "mt", iso=0, eps=np.float32(1e-3))[0].dtype == np.float32
    assert oracle.isosurface(s.astype(np.float64), "mc", iso=np.float32(0))[0].dtype == np.float64


# ---- closed forms on single cubes ---------------------------------------------

def cube_field(case, corners):
    f = np.ones((2, 2, 2))
    for c, (dx, dy, dz) in enumerate(corners):
        if (case >> c) & 1:
            f[dx, dy, dz] = -1.0
    return f


def test_mc_single_cube_all_cases(oracle):
    """Each vertex sits at the midpoint of a crossing cube edge (values +-1,
    iso 0 => mu = 1/2), every crossing edge has exactly one vertex, faces
    index 1..nv, and no face is degenerate."""
    cube_edges = [(a, b) for a, b in itertools.combinations(range(8), 2)
                  if sum(abs(p - q) for p, q in zip(MC_CORNERS[a], MC_CORNERS[b])) == 1]
    assert len(cube_edges) == 12
    for case in range(256):
        V, F = oracle.isosurface(cube_field(case, MC_CORNERS), "mc", iso=0, X=(0, 1), Y=(0, 1),
                                 Z=(0, 1))
        crossing = {tuple(np.add(MC_CORNERS[a], MC_CORNERS[b]) / 2) for a, b in cube_edges
                    if ((case >> a) & 1) != ((case >> b) & 1)}
        if case in (0, 255):
            assert len(V) == 0 and len(F) == 0
            continue
        got = [tuple(v) for v in V]
        assert len(got) == len(set(got)) == len(crossing)
        assert set(got) == crossing
        assert F.min() >= 1 and F.max() <= len(V)
        assert all(len(set(f)) == 3 for f in F.tolist())
        assert 1 <= len(F) <= 5


def test_mt_single_cube_all_cases(oracle):
    """Vertices lie on crossing edges of the 6-tetrahedron split (12 cube
    edges, 6 face diagonals, 1 body diagonal), each such edge gets one
    vertex, and vertex positions are clamped away from corners by eps."""
    diag = {(0, 2), (0, 7), (0, 5), (4, 6), (1, 6), (3, 6), (0, 6)}  # MT corner ids
    edges = [(a, b) for a, b in itertools.combinations(range(8), 2)
             if sum(abs(p - q) for p, q in zip(MT_CORNERS[a], MT_CORNERS[b])) == 1
             or (a, b) in diag]
    assert len(edges) == 19
    for case in range(256):
        V, F = oracle.isosurface(cube_field(case, MT_CORNERS), "mt", iso=0, X=(0, 1), Y=(0, 1),
                                 Z=(0, 1))
        crossing = {tuple(np.add(MT_CORNERS[a], MT_CORNERS[b]) / 2) for a, b in edges
                    if ((case >> a) & 1) != ((case >> b) & 1)}
        if case in (0, 255):
            assert len(V) == 0
            continue
        got = {tuple(np.round(v, 12)) for v in V}
        assert got == crossing
        assert sorted(np.unique(F).tolist()) == list(range(1, len(V) + 1))


# ---- closed-surface invariants ------------------------------------------------

def torus(n, R=0.6, r=0.25):
    x, y, z = grid(n)
    q = np.sqrt(x * x + y * y) - R
    return np.sqrt(q * q + z * z) - r


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_mt_closed_sphere_and_torus(oracle, dtype):
    for f, chi in ((sphere(33, 0.6, dtype), 2), (torus(41).astype(dtype), 0)):
        V, F = oracle.isosurface(f, "mt", iso=0)
        assert euler(F) == chi
        assert directed_edges_paired(F)
        assert sorted(np.unique(F).tolist()) == list(range(1, len(V) + 1))


def test_sn_closed_sphere_and_torus(oracle):
    for f, chi in ((sphere(33, 0.6), 2), (torus(41), 0)):
        V, F = oracle.isosurface(f, "sn", iso=0)
        assert euler(F) == chi
        assert directed_edges_paired(F)


def test_volume_and_orientation(oracle):
    f = sphere(64, 0.6, np.float32)
    true = 4 / 3 * np.pi * 0.6 ** 3
    for m in ("mc", "mt", "sn"):
        V, F = oracle.isosurface(f, m, iso=0)
        vol = signed_volume(V, F)
        assert vol > 0, m  # outward normals, as MC's tables give
        assert abs(vol - true) / true < 0.01, m


def test_mc_faces_stay_in_their_cube(oracle):
    # process_mc_voxel! emits faces that index only the current cube's new
    # vertices (src/marching_cubes.jl:69-90)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((9, 8, 7))
    V, F = oracle.isosurface(f, "mc", iso=0.1)
    lo = F.min(axis=1)
    assert np.all(np.diff(lo) >= 0)
    assert np.all(F.max(axis=1) - lo < 12)
    assert F.max() == len(V)


# ---- closed forms of the counts and the numbering ---------------------------

def below(f, iso):
    """The classification bit of every sample: `v < iso`, NaN gives false
    (src/common.jl:10-20)."""
    with np.errstate(invalid="ignore"):
        return np.asarray(f) < iso


def kuhn_crossing_edges(b):
    """Crossing edges of the 6-tetrahedron split of every voxel, on the
    global lattice: the 3 axis edges, the 3 face diagonals that run from a
    face's low corner to its high corner, and the body diagonal (the split
    of src/lut/mt.jl shares corners (0,0,0) and (1,1,1) in every
    tetrahedron).  Each edge is counted once."""
    n = 0
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]:
        lo = b[:b.shape[0] - d[0], :b.shape[1] - d[1], :b.shape[2] - d[2]]
        hi = b[d[0]:, d[1]:, d[2]:]
        n += int((lo != hi).sum())
    return n


def kuhn_triangles(b):
    """Triangles of the 6 tetrahedra (0,0,0), e_a, e_a+e_b, (1,1,1) of every
    voxel: one when 1 or 3 of the 4 corners are below, two when 2 are."""
    c0 = b[:-1, :-1, :-1].astype(np.int8)
    c7 = b[1:, 1:, 1:].astype(np.int8)
    unit = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
    nx, ny, nz = b.shape

    def at(d):
        return b[d[0]:nx - 1 + d[0], d[1]:ny - 1 + d[1], d[2]:nz - 1 + d[2]].astype(np.int8)

    total = 0
    for a, c in itertools.permutations(range(3), 2):
        ea = unit[a]
        eab = tuple(p + q for p, q in zip(unit[a], unit[c]))
        k = c0 + at(ea) + at(eab) + c7
        total += int(((k == 1) | (k == 3)).sum()) + 2 * int((k == 2).sum())
    return total


def first_appearance_order(F):
    """Vertex indices in the order they first appear in the flattened faces."""
    flat = F.reshape(-1)
    _, first = np.unique(flat, return_index=True)
    return flat[np.sort(first)]


MT_FIELDS = [
    ("noise", lambda: np.random.default_rng(11).standard_normal((9, 8, 7)), 0.0),
    ("noise-iso", lambda: np.random.default_rng(12).standard_normal((6, 10, 5)).astype(np.float32),
     0.3),
    ("sphere", lambda: sphere(17, 0.55), 0.0),
    ("torus", lambda: torus(15), 0.0),
    ("thin", lambda: np.random.default_rng(13).standard_normal((2, 7, 9)), -0.2),
]


@pytest.mark.parametrize("name,make,iso", MT_FIELDS, ids=[m[0] for m in MT_FIELDS])
def test_mt_counts_closed_form(oracle, name, make, iso):
    # getVertId creates one vertex per distinct crossing tetrahedron edge
    # (vertId names an edge the same from every voxel,
    # src/marching_tetrahedra.jl:29-32,67-84) and tetTri gives 1 or 2
    # triangles per cut tetrahedron.  A slip in the Dict key (two edges one
    # key, or one edge two keys) changes V.
    f = make()
    V, F = oracle.isosurface(f, "mt", iso=iso)
    b = below(f, iso)
    assert len(V) == kuhn_crossing_edges(b) > 0
    assert len(F) == kuhn_triangles(b)


@pytest.mark.parametrize("name,make,iso", MT_FIELDS, ids=[m[0] for m in MT_FIELDS])
def test_mt_numbering_is_first_request(oracle, name, make, iso):
    # A vertex is numbered length(vtsAry)+1 when it is first requested, and
    # every request is an argument of a face pushed at once, left to right
    # (src/marching_tetrahedra.jl:67-84,117-125).  So the vertex indices, in
    # the order they first appear in the flattened faces, are exactly 1..V.
    # Numbering a voxel's vertices before its faces, or evaluating a face's
    # arguments in another order, breaks this.
    V, F = oracle.isosurface(make(), "mt", iso=iso)
    assert np.array_equal(first_appearance_order(F), np.arange(1, len(V) + 1))
    # the check has teeth: the same faces with reversed arguments fail it
    assert not np.array_equal(first_appearance_order(F[:, ::-1].copy()),
                              np.arange(1, len(V) + 1))


def test_mc_numbering_is_first_appearance(oracle):
    # process_mc_voxel! pushes a cube's vertices in _mc_verts order and its
    # first face is (f+3, f+2, f+1) (src/marching_cubes.jl:69-90), so the
    # vertices appear in order cube by cube: reading each face backwards,
    # first appearances are 1..V
    f = np.random.default_rng(14).standard_normal((8, 9, 7))
    V, F = oracle.isosurface(f, "mc", iso=0.1)
    assert np.array_equal(first_appearance_order(F[:, ::-1].copy()), np.arange(1, len(V) + 1))


SN_FIELDS = [
    ("noise", lambda: np.random.default_rng(4).standard_normal((12, 11, 10)), 0.0),
    ("noise-f32", lambda: np.random.default_rng(5).standard_normal((7, 9, 8)).astype(np.float32),
     0.25),
    ("sphere", lambda: sphere(19, 0.6), 0.0),
    ("thin", lambda: np.random.default_rng(6).standard_normal((2, 6, 5)), 0.0),
]


@pytest.mark.parametrize("name,make,iso", SN_FIELDS, ids=[m[0] for m in SN_FIELDS])
def test_sn_spec_invariants(oracle, name, make, iso):
    # The surface-nets spec (one vertex per active cell, numbered in the
    # reference's cell order, placed in its cell; one quad per crossing grid
    # edge whose four cells exist, emitted by the cell at the edge's low
    # corner, as two triangles that share that cell's vertex), checked
    # against counts and positions computed here from the samples alone.
    f = make()
    nx, ny, nz = f.shape
    X, Y, Z = (0.0, float(nx - 1)), (-2.0, 2.0 * (ny - 1) - 2.0), (1.0, 1.0 + 0.5 * (nz - 1))
    V, F = oracle.isosurface(f, "sn", iso=iso, X=X, Y=Y, Z=Z)
    b = below(f, iso)
    corners = [b[dx:nx - 1 + dx, dy:ny - 1 + dy, dz:nz - 1 + dz]
               for dx, dy, dz in MC_CORNERS]
    s = sum(c.astype(np.int8) for c in corners)
    active = (s > 0) & (s < 8)
    cells = np.argwhere(active)  # row-major over (x, y, z): z fastest, the reference's order
    assert len(V) == len(cells)
    # each vertex lies in its cell (up to the rounding of a mean)
    lo = np.array([X[0], Y[0], Z[0]]) + cells * np.array(
        [(X[1] - X[0]) / (nx - 1), (Y[1] - Y[0]) / (ny - 1), (Z[1] - Z[0]) / (nz - 1)])
    hi = lo + np.array([(X[1] - X[0]) / (nx - 1), (Y[1] - Y[0]) / (ny - 1),
                        (Z[1] - Z[0]) / (nz - 1)])
    tol = 1e-5 if V.dtype == np.float32 else 1e-12
    assert np.all(V >= lo - tol) and np.all(V <= hi + tol)
    # quads: crossing edges with all four cells present
    q = int((b[:-1, 1:-1, 1:-1] != b[1:, 1:-1, 1:-1]).sum()
            + (b[1:-1, :-1, 1:-1] != b[1:-1, 1:, 1:-1]).sum()
            + (b[1:-1, 1:-1, :-1] != b[1:-1, 1:-1, 1:]).sum())
    assert len(F) == 2 * q
    if q:
        # every face leads with the emitting cell's vertex, the largest index
        # of the face, in cell order; the two triangles of a quad share it
        # and the quad's diagonal
        assert np.all(F[:, 0] > F[:, 1]) and np.all(F[:, 0] > F[:, 2])
        assert np.all(np.diff(F[:, 0]) >= 0)
        assert np.array_equal(F[0::2, 0], F[1::2, 0]) and np.array_equal(F[0::2, 2], F[1::2, 1])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_xslab_oracle_concatenates_to_the_whole(oracle, dtype):
    # the x-slab form of the Marching Cubes oracle, used to check meshes
    # too large for the whole-field oracle: slabs at the start, middle and
    # end concatenate (faces shifted) to the whole mesh
    f = np.random.default_rng(21).standard_normal((13, 9, 8)).astype(dtype)
    nx = f.shape[0]
    V, F = oracle.isosurface(f, "mc", iso=0.05)
    parts, nv = [], 0
    for xlo, xhi in ((1, 3), (4, 4), (5, 9), (10, 12)):
        slab = f[xlo - 1:xhi + 1]
        Vs, Fs = oracle.isosurface_xslab(slab, nx, xlo - 1, xlo, xhi, iso=0.05)
        parts.append((Vs, Fs + nv))
        nv += len(Vs)
    assert np.array_equal(np.concatenate([p[0] for p in parts]), V)
    assert np.array_equal(np.concatenate([p[1] for p in parts]), F)


def test_empty_and_degenerate(oracle):
    for shape in ((0, 0, 0), (2, 2, 0), (3, 1, 3)):
        f = np.zeros(shape)
        if 1 in shape:
            with pytest.raises(ValueError, match="endpoints differ"):
                oracle.isosurface(f, "mc")
            V, F = oracle.isosurface(f, "mc", Y=(0, 0))
        else:
            V, F = oracle.isosurface(f, "mc")
        assert len(V) == 0 and len(F) == 0
    V, F = oracle.isosurface(np.ones((5, 5, 5)), "mt", iso=0)
    assert len(V) == 0


@pytest.mark.parametrize("dtype,shape,threads", [(np.float64, (23, 9, 11), 4),
                                                  (np.float32, (7, 12, 10), 16),
                                                  (np.float64, (2, 5, 5), 3)])
def test_mc_threaded_equals_serial(oracle, dtype, shape, threads):
    # the multi-threaded CPU baseline (x-slabs on threads) returns the
    # single-threaded result exactly, including more threads than x-cells
    rng = np.random.default_rng(5)
    f = rng.standard_normal(shape).astype(dtype)
    V1, F1 = oracle.isosurface(f, "mc", iso=0.05)
    Vt, Ft = oracle.isosurface(f, "mc", iso=0.05, threads=threads)
    assert len(F1) > 0
    assert np.array_equal(V1, Vt) and np.array_equal(F1, Ft)
