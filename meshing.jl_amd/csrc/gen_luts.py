#!/usr/bin/env python3
"""Generate ``luts.h``: the lookup tables of the isosurface kernels, in the
packed, 0-based layout the HIP kernels read.

The source of truth is the reference's table file text (Meshing.jl v0.7.0,
``src/lut/mc.jl`` and ``src/lut/mt.jl``, MIT licence, (c) 2015 Steve Kelly,
Tracy Wadleigh).  The tables are data that must be identical to the
reference's for the output to be identical; everything below the parse step
*derives* new tables (per-case vertex/triangle lists for Marching Cubes,
per-case face lists, vertex-ownership and first-appearance ranks for
Marching Tetrahedra) that the reference computes on the fly.

Run once in a checkout where the reference is mounted::

    python3 meshing.jl_amd/csrc/gen_luts.py /root/reference > meshing.jl_amd/csrc/luts.h

The generated header is committed; the GPU build never needs the reference.
"""
import ast
import re
import sys
from pathlib import Path


def parse_julia_consts(text):
    """Return {name: tuple} for every ``const name = (...)`` in ``text``.

    Julia tuple literals of integer/hex literals are also Python literals once
    comments are stripped, so ``ast.literal_eval`` reads them directly."""
    out = {}
    for m in re.finditer(r"const\s+(\w+)\s*=\s*\(", text):
        name, i, depth = m.group(1), m.end() - 1, 0
        for j in range(i, len(text)):
            if text[j] == "(":
                depth += 1
            elif text[j] == ")":
                depth -= 1
                if depth == 0:
                    break
        body = re.sub(r"#[^\n]*", "", text[i:j + 1])
        out[name] = ast.literal_eval(body)
    return out


# Marching Cubes corner order of src/marching_cubes.jl:42-49, as (dx,dy,dz).
MC_CORNERS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
              (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def mc_tables(t):
    verts, eqmap, conn, edges = (t["_mc_verts"], t["_mc_eq_mapping"],
                                 t["_mc_connectivity"], t["_mc_edge_list"])
    assert len(verts) == 254 and len(eqmap) == 254 and len(edges) == 12
    case_verts = [0] * 256   # bits 0-3 nv, then 12 nibbles: 0-based edge ids
    case_tris = [0] * 256    # bits 0-2 nt, then 15 nibbles: 0-based local ids
    for c in range(1, 255):
        vl = [v - 1 for v in verts[c - 1] if v != 0]
        # vertex list is a prefix: the reference stops at the first zero
        # (src/marching_cubes.jl:74-76)
        first_zero = list(verts[c - 1]).index(0) if 0 in verts[c - 1] else 12
        assert first_zero == len(vl)
        word = len(vl)
        for q, e in enumerate(vl):
            word |= e << (4 + 4 * q)
        case_verts[c] = word
        # faces in emission order (src/marching_cubes.jl:82-90): the first
        # triangle is (3,2,1), then each nonzero triple o[i..i+2] of the
        # connectivity row is emitted reversed.
        tris = [(2, 1, 0)]
        o = conn[eqmap[c - 1] - 1]
        for i in (0, 3, 6, 9):
            if o[i] == 0:
                break
            tris.append((o[i + 2] - 1, o[i + 1] - 1, o[i] - 1))
        word = len(tris)
        for k, tri in enumerate(tris):
            for m, lv in enumerate(tri):
                assert 0 <= lv < len(vl), (c, tri)
                word |= lv << (3 + 4 * (3 * k + m))
        case_tris[c] = word
    edge_words = [(a - 1) | ((b - 1) << 4) for a, b in edges]
    return case_verts, case_tris, edge_words


def mt_tables(t):
    pos = t["voxCrnrPos"]
    ecrn = t["voxEdgeCrnrs"]
    edir = t["voxEdgeDir"]
    eix = t["voxEdgeIx"]
    subtets = t["subTets"]
    masks = t["subTetsMask"]
    tetedge = t["tetEdgeCrnrs"]
    tettri = t["tetTri"]
    assert len(ecrn) == 19 and len(edir) == 19

    def tet_ix(i, case):          # src/marching_tetrahedra.jl:11-18
        v1, v2 = masks[i]
        ix = (0x01 & case) | ((0x40 & case) >> 3)
        if v1 & case:
            ix |= 0x02
        if v2 & case:
            ix |= 0x04
        return ix

    def vox_edge(i, te):          # src/marching_tetrahedra.jl:91-95
        s = subtets[i][tetedge[te - 1][0] - 1]
        g = subtets[i][tetedge[te - 1][1] - 1]
        return eix[s - 1][g - 1]

    # faces per case in emission order: tets 1..6, then tetTri entries
    # (src/marching_tetrahedra.jl:110-126); 0-based voxel-edge ids.
    nf = [0] * 256
    faces = [[0] * 36 for _ in range(256)]
    for case in range(256):
        if case in (0, 255):
            continue
        fl = []
        for i in range(6):
            tix = tet_ix(i, case)
            if tix in (0, 15):
                continue
            e = tettri[tix - 1]
            fl.append(tuple(vox_edge(i, e[k]) - 1 for k in range(3)))
            if e[3] == 0:
                continue
            fl.append(tuple(vox_edge(i, e[k]) - 1 for k in range(3, 6)))
        nf[case] = len(fl)
        for k, f in enumerate(fl):
            faces[case][3 * k:3 * k + 3] = list(f)

    end_corner = [pos[ecrn[e][1] - 1] for e in range(19)]
    start_corner = [pos[ecrn[e][0] - 1] for e in range(19)]

    # Ownership: a vertex is created by the first voxel (in i,j,k order)
    # that requests its edge.  Every voxel containing a crossing edge
    # requests it, so the owner of an edge ending at global grid point q is
    # voxel max(q-1, first) per axis.  Seen from a voxel, edge e is owned
    # iff along each axis its end corner is at local 1, or at local 0 with
    # the voxel on the first layer of that axis.  class bit a = "voxel is on
    # the first layer along axis a".
    def owned(e, cls):
        for a in range(3):
            if end_corner[e][a] == 0 and not (cls >> a) & 1:
                return False
        return True

    nown = [[0] * 256 for _ in range(8)]
    own_edges = [[[255] * 20 for _ in range(256)] for _ in range(8)]
    rank = [[[255] * 19 for _ in range(256)] for _ in range(8)]
    for cls in range(8):
        for case in range(256):
            seen = []
            for k in range(nf[case]):
                for m in range(3):
                    e = faces[case][3 * k + m]
                    if owned(e, cls) and e not in seen:
                        seen.append(e)
            nown[cls][case] = len(seen)
            for q, e in enumerate(seen):
                own_edges[cls][case][q] = e
                rank[cls][case][e] = q

    # Shifting an edge by delta (dx,dy,dz) in {0,1}^3 into the voxel at
    # -delta: the same grid edge seen from the neighbour.
    corner_of = {tuple(p): c for c, p in enumerate(pos)}
    shift = [[255] * 8 for _ in range(19)]
    for e in range(19):
        for d in range(8):
            dv = (d & 1, (d >> 1) & 1, (d >> 2) & 1)
            a = tuple(start_corner[e][k] + dv[k] for k in range(3))
            b = tuple(end_corner[e][k] + dv[k] for k in range(3))
            if a in corner_of and b in corner_of:
                ca, cb = corner_of[a], corner_of[b]
                ee = eix[ca][cb] - 1
                assert ee >= 0 and ecrn[ee] == (ca + 1, cb + 1)
                shift[e][d] = ee

    corner_words = [p[0] | (p[1] << 1) | (p[2] << 2) for p in pos]
    edge_words = [(a - 1) | ((b - 1) << 4) for a, b in ecrn]
    return dict(nf=nf, faces=faces, nown=nown, own_edges=own_edges,
                rank=rank, shift=shift, corner_words=corner_words,
                edge_words=edge_words, edir=list(edir))


def c_array(ctype, name, values, per_line=12, fmt="{}"):
    flat = values
    lines = []
    for i in range(0, len(flat), per_line):
        lines.append("    " + ", ".join(fmt.format(v) for v in flat[i:i + per_line]) + ",")
    return f"MSH_LUT {ctype} {name}[{len(flat)}] = {{\n" + "\n".join(lines) + "\n};\n"


def main(ref_root):
    ref = Path(ref_root)
    mc = parse_julia_consts((ref / "src/lut/mc.jl").read_text())
    mt = parse_julia_consts((ref / "src/lut/mt.jl").read_text())
    case_verts, case_tris, mc_edges = mc_tables(mc)
    T = mt_tables(mt)

    flat = lambda xss: [x for xs in xss for x in xs]
    out = []
    out.append("// Generated by gen_luts.py from the table text of Meshing.jl v0.7.0\n"
               "// (src/lut/mc.jl, src/lut/mt.jl; MIT licence, (c) 2015 Steve Kelly,\n"
               "// Tracy Wadleigh).  Do not edit by hand.\n"
               "#pragma once\n#include <stdint.h>\n"
               "#ifndef MSH_LUT\n#define MSH_LUT static const\n#endif\n")
    out.append("// Marching Cubes (0-based).  MC_EDGE[e] = corner_a | corner_b << 4, the\n"
               "// interpolation orientation of src/lut/mc.jl:606-608.\n")
    out.append(c_array("uint8_t", "MC_EDGE", mc_edges))
    out.append("// MC_CASE_VERTS[case]: bits 0-3 vertex count, then one nibble per vertex\n"
               "// holding its edge id, in the order vertices are pushed.\n")
    out.append(c_array("uint64_t", "MC_CASE_VERTS", case_verts, 4, "0x{:016x}ull"))
    out.append("// MC_CASE_TRIS[case]: bits 0-2 triangle count, then 3 nibbles per\n"
               "// triangle holding local vertex ids in the order faces are pushed.\n")
    out.append(c_array("uint64_t", "MC_CASE_TRIS", case_tris, 4, "0x{:016x}ull"))
    out.append("// Marching Tetrahedra (0-based).  MT_CORNER[c] = dx | dy<<1 | dz<<2;\n"
               "// MT_EDGE[e] = start corner | end corner << 4; MT_EDGE_DIR = direction code.\n")
    out.append(c_array("uint8_t", "MT_CORNER", T["corner_words"]))
    out.append(c_array("uint8_t", "MT_EDGE", T["edge_words"]))
    out.append(c_array("uint8_t", "MT_EDGE_DIR", T["edir"]))
    out.append("// MT_NF[case]: faces per case; MT_FACE[case*36 + 3k + m]: voxel edge of\n"
               "// corner m of face k, in emission order.\n")
    out.append(c_array("uint8_t", "MT_NF", T["nf"], 16))
    out.append(c_array("uint8_t", "MT_FACE", flat(T["faces"]), 18))
    out.append("// Vertex ownership.  cls = (first x layer) | (first y layer)<<1 |\n"
               "// (first z layer)<<2.  MT_NOWN[cls*256+case]: vertices the voxel creates;\n"
               "// MT_OWN_EDGES[(cls*256+case)*20 + q]: edge of its q-th new vertex;\n"
               "// MT_RANK[(cls*256+case)*19 + e]: q of edge e (255: not created here).\n")
    out.append(c_array("uint8_t", "MT_NOWN", flat(T["nown"]), 16))
    out.append(c_array("uint8_t", "MT_OWN_EDGES", flat(flat(T["own_edges"])), 20))
    out.append(c_array("uint8_t", "MT_RANK", flat(flat(T["rank"])), 19))
    out.append("// MT_SHIFT[e*8 + d]: the voxel edge, seen from the neighbour at -d\n"
               "// (d = dx | dy<<1 | dz<<2), that is the same grid edge as e (255: none).\n")
    out.append(c_array("uint8_t", "MT_SHIFT", flat(T["shift"]), 8))
    sys.stdout.write("\n".join(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
