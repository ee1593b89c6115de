"""CPU checks of the kernels' lookup tables (meshing.jl_amd/csrc/luts.h).

The kernels never walk the reference's per-cube logic; they read derived
tables: per-case vertex/triangle lists for Marching Cubes and, for Marching
Tetrahedra, which edges a voxel creates vertices for (ownership), in which
order (rank), and how to find the owner of an edge (shift).  Here a numpy
model that uses only those tables (the same indexing the kernels do) must
reproduce the oracle's faces exactly, and the tables must agree with the
cube geometry.
"""
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
LUTS = ROOT / "meshing.jl_amd" / "csrc" / "luts.h"


def parse_luts():
    text = LUTS.read_text()
    out = {}
    for m in re.finditer(r"MSH_LUT\s+(\w+)\s+(\w+)\[(\d+)\]\s*=\s*\{([^}]*)\}", text):
        ctype, name, n, body = m.groups()
        vals = [int(v.strip().rstrip("ull"), 0) for v in body.replace("\n", " ").split(",")
                if v.strip()]
        assert len(vals) == int(n), name
        out[name] = np.array(vals, dtype=np.uint64 if ctype == "uint64_t" else np.int64)
    return out


T = parse_luts()
MC_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def cases(f, iso, corners):
    nx, ny, nz = f.shape
    c = np.zeros((nx - 1, ny - 1, nz - 1), dtype=np.int64)
    for b, (dx, dy, dz) in enumerate(corners):
        c |= (f[dx:nx - 1 + dx, dy:ny - 1 + dy, dz:nz - 1 + dz] < iso).astype(np.int64) << b
    return c


def mc_model_faces(f, iso):
    cs = cases(f, iso, MC_CORNERS).ravel()  # C order over (i, j, k) = emission order
    nv = (T["MC_CASE_VERTS"][cs] & np.uint64(0xF)).astype(np.int64)
    voff = np.cumsum(nv) - nv
    out = []
    for c, base in zip(cs, voff):
        tw = int(T["MC_CASE_TRIS"][c])
        for t in range(tw & 7):
            w = tw >> (3 + 12 * t)
            out.append((base + 1 + (w & 15), base + 1 + ((w >> 4) & 15), base + 1 + ((w >> 8) & 15)))
    return int(nv.sum()), np.array(out, dtype=np.int64).reshape(-1, 3)


def mt_model_faces(f, iso):
    corners = [((w & 1), (w >> 1) & 1, (w >> 2) & 1) for w in T["MT_CORNER"].tolist()]
    cs = cases(f, iso, corners)
    ni, nj, nk = cs.shape
    ii, jj, kk = np.meshgrid(np.arange(ni), np.arange(nj), np.arange(nk), indexing="ij")
    cls = (ii == 0) | ((jj == 0) << 1) | ((kk == 0) << 2)
    nown = T["MT_NOWN"][cls * 256 + cs]
    voff = (np.cumsum(nown.ravel()) - nown.ravel()).reshape(cs.shape)
    end = [corners[int(w) >> 4] for w in T["MT_EDGE"].tolist()]
    out = []
    for i in range(ni):
        for j in range(nj):
            for k in range(nk):
                c = cs[i, j, k]
                for t in range(int(T["MT_NF"][c])):
                    tri = []
                    for m in range(3):
                        e = int(T["MT_FACE"][c * 36 + 3 * t + m])
                        d = [1 if end[e][a] == 0 and (i, j, k)[a] > 0 else 0 for a in range(3)]
                        e2 = int(T["MT_SHIFT"][e * 8 + d[0] + 2 * d[1] + 4 * d[2]])
                        assert e2 != 255
                        oi, oj, ok = i - d[0], j - d[1], k - d[2]
                        oc = cs[oi, oj, ok]
                        rank = int(T["MT_RANK"][(cls[oi, oj, ok] * 256 + oc) * 19 + e2])
                        assert rank != 255
                        tri.append(voff[oi, oj, ok] + rank + 1)
                    out.append(tri)
    return int(nown.sum()), np.array(out, dtype=np.int64).reshape(-1, 3)


def fields():
    rng = np.random.default_rng(12345)
    yield rng.standard_normal((6, 5, 7)), 0.0
    yield rng.standard_normal((4, 9, 3)), 0.3
    yield rng.uniform(-1, 1, (8, 8, 8)), -0.2
    r = np.linspace(-1, 1, 14)
    x, y, z = np.meshgrid(r, r, r, indexing="ij")
    yield np.sqrt(x * x + y * y + z * z) - 0.7, 0.0


@pytest.mark.parametrize("k", range(4))
def test_mc_tables_reproduce_oracle_faces(oracle, k):
    f, iso = list(fields())[k]
    V, F = oracle.isosurface(f, "mc", iso=iso)
    nv, Fm = mc_model_faces(f, iso)
    assert nv == len(V)
    assert np.array_equal(Fm, F)


@pytest.mark.parametrize("k", range(4))
def test_mt_tables_reproduce_oracle_faces(oracle, k):
    """The owner rule + first-appearance ranks number vertices exactly as the
    reference's Dict does (src/marching_tetrahedra.jl:67-84)."""
    f, iso = list(fields())[k]
    V, F = oracle.isosurface(f, "mt", iso=iso)
    nv, Fm = mt_model_faces(f, iso)
    assert nv == len(V)
    assert np.array_equal(Fm, F)


def test_mc_vertex_edges_are_crossing_edges():
    edges = [(int(w) & 15, int(w) >> 4) for w in T["MC_EDGE"].tolist()]
    for a, b in edges:
        assert sum(abs(p - q) for p, q in zip(MC_CORNERS[a], MC_CORNERS[b])) == 1
    hist_t, hist_v = {}, {}
    for c in range(256):
        vw, tw = int(T["MC_CASE_VERTS"][c]), int(T["MC_CASE_TRIS"][c])
        nv, nt = vw & 15, tw & 7
        vl = [(vw >> (4 + 4 * q)) & 15 for q in range(nv)]
        crossing = {e for e, (a, b) in enumerate(edges) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert set(vl) == crossing and len(vl) == len(crossing)
        if c in (0, 255):
            assert nv == 0 and nt == 0
            continue
        hist_t[nt] = hist_t.get(nt, 0) + 1
        hist_v[nv] = hist_v.get(nv, 0) + 1
        used = set()
        for t in range(nt):
            w = tw >> (3 + 12 * t)
            tri = (w & 15, (w >> 4) & 15, (w >> 8) & 15)
            assert len(set(tri)) == 3 and max(tri) < nv
            used |= set(tri)
        assert used == set(range(nv))
        # the first emitted triangle is (3, 2, 1) (src/marching_cubes.jl:82)
        assert (tw >> 3) & 0xFFF == (2 | (1 << 4) | (0 << 8))
    # distribution over the 254 cases (classic table)
    assert hist_t == {1: 16, 2: 50, 3: 80, 4: 76, 5: 32}
    assert hist_v == {3: 16, 4: 30, 5: 48, 6: 64, 7: 48, 8: 30, 9: 16, 12: 2}


def test_mt_faces_use_exactly_crossing_edges():
    corners = [((w & 1), (w >> 1) & 1, (w >> 2) & 1) for w in T["MT_CORNER"].tolist()]
    edges = [(int(w) & 15, int(w) >> 4) for w in T["MT_EDGE"].tolist()]
    for a, b in edges:  # every edge runs from the lower to the upper corner
        assert all(corners[a][k] <= corners[b][k] for k in range(3))
    for c in range(256):
        used = {int(T["MT_FACE"][c * 36 + q]) for q in range(3 * int(T["MT_NF"][c]))}
        crossing = {e for e, (a, b) in enumerate(edges) if ((c >> a) & 1) != ((c >> b) & 1)}
        assert used == crossing
        # ranks are a permutation of the owned crossing edges
        for cls in range(8):
            n = int(T["MT_NOWN"][cls * 256 + c])
            own = [int(T["MT_OWN_EDGES"][(cls * 256 + c) * 20 + q]) for q in range(n)]
            for q, e in enumerate(own):
                assert int(T["MT_RANK"][(cls * 256 + c) * 19 + e]) == q
            if cls == 7:  # the very first voxel owns every crossing edge
                assert set(own) == crossing


@pytest.mark.skipif(not Path("/root/reference/src/lut/mc.jl").exists(),
                    reason="reference tables not mounted")
def test_generated_header_is_current():
    gen = ROOT / "meshing.jl_amd" / "csrc" / "gen_luts.py"
    out = subprocess.run([sys.executable, str(gen), "/root/reference"], capture_output=True,
                         text=True, check=True).stdout
    assert out == LUTS.read_text()
