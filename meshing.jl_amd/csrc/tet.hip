// tet.hip -- the Marching Tetrahedra emit kernel (src/marching_tetrahedra.jl:105-163)
// and the Naive Surface Nets emit kernel (this library's definition, see
// oracle/cpu_ref.cpp sn_isosurface).
//
// Same pipeline as the Marching Cubes emit kernel (lean.hip): the sample
// brick of a tile arrives in LDS by DMA, one ballot per sample row gives the
// below-iso masks, lanes take (row, z) cells with a segmented DPP scan for
// the in-row output prefixes, active cells go to a per-wave list, and the
// list is emitted slot-parallel.
//
// What is particular to tetrahedra: a face refers to vertices owned by the
// cells at -x, -y and -z (the owner of a grid edge is the first voxel in
// traversal order that contains it, which is what the reference's Dict
// gives, src/marching_tetrahedra.jl:67-84).  Each wave therefore classifies
// one column of cells at -x and one row at -y next to the cells it emits,
// and walks its rows so that every owner is classified (case and in-row
// vertex prefix in LDS) before any face that names it is written; the wave
// can emit its list whenever it fills, without waiting for the others.
// Surface nets have the same structure: a quad joins the vertices of the
// four cells around a crossing grid edge, at -u, -w and -u-w.
#include <stdlib.h>

#include "device.h"

#define MSH_LUT static __constant__
#include "luts.h"

namespace msh {
namespace {

constexpr int kTetWX = 14;               // cells a wave emits along x
constexpr int kTetTX = kWaves * kTetWX;  // 56 per block
constexpr int kTetZC = 16;               // z chunk

// Row r of a wave: r = yr * 16 + xr.  yr 0 is the row at -y, yr 1 the row
// the block emits; xr 0 is the column at -x, 1..14 the wave's cells, 15 is
// padding.  Brick sample columns start 4 cells before the tile so that rows
// stay 16-byte aligned: cell i of the tile sits at brick column i + 4.
template <int METHOD, typename T, typename M, typename A, typename Out, bool VEC, int CXM>
__device__ __forceinline__ void emit_halo_tile(const VolArgs& a, int64_t t) {
  static_assert(METHOD == MT || METHOD == SN, "tetrahedra or surface nets");
  constexpr int ZC = kTetZC, R = kWave / ZC;
  constexpr int SX = 64, SY = 3, NP = ZC + 2;  // planes zc-1 .. zc+ZC
  constexpr int DW = (int)(sizeof(T) / 4), RD = SX * DW;
  constexpr int WR = 32, NSTEP = WR / R;
  constexpr int CAP = 128;
  constexpr int VW = (CAP * 13 + 63) / 64, FW = (CAP * 12 + 63) / 64;
  constexpr int K = METHOD == MT ? 1 : 0;
  struct Smem {
    uint32_t smp[2][NP * SY * RD];
    double cz[2][NP];
    double cx[SX], cy[SY];
    uint64_t mask[NP * SY];        // [plane][y], bit x
    // tetrahedra only (size 1 for surface nets)
    uint8_t nown[K ? 8 * 256 : 1];       // vertices created, by class and (natural) case
    uint8_t nf[K ? 256 : 1];             // faces, by case
    uint16_t face[K ? 256 * 12 : 1];     // edges of each face, 5 bits each
    uint8_t rank0[K ? 256 * 19 : 1];     // rank of each edge among a class-0 voxel's vertices
    uint8_t own0[K ? 256 * 8 : 1];       // edges of a class-0 voxel's vertices, in order
    uint8_t shift[K ? 152 : 1];          // MT_SHIFT
    uint8_t ccase[kWaves][WR][K ? ZC : 1];
    uint8_t pcase[kWaves][K ? WR : 1];   // case of the cell just below the chunk
    uint32_t edge[19];                   // per edge: corner offsets and axis bits
    uint16_t cpv[kWaves][WR][ZC];        // in-row vertex prefix of each cell of the chunk
    int64_t rbv[2][kWaves][WR], rbf[2][kWaves][WR];
    // cells with vertices / with faces (a cell may have one and not the
    // other), each with its first slot
    uint64_t vlist[kWaves][CAP], flist[kWaves][CAP];
    uint16_t lpv[kWaves][CAP], lpf[kWaves][CAP];
    uint64_t vbits[kWaves][VW], fbits[kWaves][FW];
  };
  __shared__ Smem sm;
  const T* __restrict__ f = (const T*)a.f;
  Out* __restrict__ verts = (Out*)a.verts;
  int64_t* __restrict__ faces = a.faces;
  const Geom g = a.g;
  const int lane = lane_id(), wv = wave_id();
  const int jb = (int)(t % a.ntj);
  const int64_t rest = t / a.ntj;
  const int ib = (int)(rest % a.nti) * kTetTX;
  const int s = (int)(rest / a.nti);
  const int zs = s * g.zseg, ze = min(zs + g.zseg, g.nz - 1);
  const T thr = (T)a.thr;
  const M iso = (M)a.iso;
  const A eps = (A)a.eps, ome = (A)a.ome;
  const int cxm = CXM >= 0 ? CXM : a.cxmask;
  const bool cx32 = cxm & 1, cy32 = (cxm >> 1) & 1, cz32 = (cxm >> 2) & 1;
  const double* zp = a.coords + g.nx + g.ny;
  const int64_t vbase = a.sz[0], totv = a.sz[1], fbase = a.sz[2], totf = a.sz[3];
  const int64_t pstride = (int64_t)g.nx * g.ny;
  const int64_t nelem = pstride * g.nz;
  const int x0 = ib - 4, y0 = jb - 1;  // brick origin (sample column 0, row 0)

  // tables, re-indexed by the natural-order case
  for (int c = threadIdx.x; METHOD == MT && c < 256; c += kBlock) {
    const uint32_t m = nat_to_mt((uint32_t)c);
#pragma unroll
    for (int cls = 0; cls < 8; ++cls) sm.nown[cls * 256 + c] = MT_NOWN[cls * 256 + m];
    const int nf = MT_NF[m];
    sm.nf[c] = (uint8_t)nf;
    for (int q = 0; q < 12; ++q) {
      uint16_t w = 0;
      if (q < nf)
        w = (uint16_t)(MT_FACE[m * 36 + 3 * q] | (MT_FACE[m * 36 + 3 * q + 1] << 5) |
                       (MT_FACE[m * 36 + 3 * q + 2] << 10));
      sm.face[c * 12 + q] = w;
    }
    for (int e = 0; e < 19; ++e) sm.rank0[c * 19 + e] = MT_RANK[m * 19 + e];
    for (int q = 0; q < 8; ++q) sm.own0[c * 8 + q] = q < 7 ? MT_OWN_EDGES[m * 20 + q] : 0;
  }
  if (threadIdx.x < (METHOD == MT ? 19 : 12)) {
    // edge e from corner ca to cb (tetrahedron edges in MT corner order, or
    // the 12 cube edges in MC corner order): brick offsets (dwords from the
    // cell's corner 0) | offset of cb << 10 | (dx,dy,dz) of ca << 20 | of
    // cb << 23 | ca << 26 | cb << 29 (corner ids, MC order, surface nets)
    const uint32_t ew = METHOD == MT ? MT_EDGE[threadIdx.x] : MC_EDGE[threadIdx.x];
    const int ca = ew & 0xF, cb = ew >> 4;
    const int xa = METHOD == MT ? mt_dx(ca) : mc_dx(ca), ya = METHOD == MT ? mt_dy(ca) : mc_dy(ca),
              za = METHOD == MT ? mt_dz(ca) : mc_dz(ca);
    const int xb = METHOD == MT ? mt_dx(cb) : mc_dx(cb), yb = METHOD == MT ? mt_dy(cb) : mc_dy(cb),
              zb = METHOD == MT ? mt_dz(cb) : mc_dz(cb);
    static_assert((SY + 1) * RD + DW < (1 << 10), "corner offsets fit 10 bits");
    const uint32_t oa = (uint32_t)((za * SY + ya) * RD + xa * DW);
    const uint32_t ob = (uint32_t)((zb * SY + yb) * RD + xb * DW);
    sm.edge[threadIdx.x] = oa | (ob << 10) | ((uint32_t)(xa | (ya << 1) | (za << 2)) << 20) |
                           ((uint32_t)(xb | (yb << 1) | (zb << 2)) << 23) |
                           (METHOD == SN ? ((uint32_t)ca << 26) | ((uint32_t)cb << 29) : 0u);
  }
  for (int q = threadIdx.x; METHOD == MT && q < 152; q += kBlock) sm.shift[q] = MT_SHIFT[q];
  for (int q = threadIdx.x; q < SX; q += kBlock)
    sm.cx[q] = a.coords[min(max(x0 + q, 0), g.nx - 1)];
  if (threadIdx.x < SY) sm.cy[threadIdx.x] = a.coords[g.nx + min(max(y0 + (int)threadIdx.x, 0), g.ny - 1)];
  bool any = false;  // some row of the brick, halo rows included, has outputs
  for (int q = threadIdx.x; q < kWaves * WR; q += kBlock) {
    const int w = q / WR, r = q % WR;
    const int xr = r & 15, yr = r >> 4;
    const int gi = ib + w * kTetWX + xr - 1, gj = jb - 1 + yr;
    int64_t bv = 0, bf = 0;
    if (xr < 15 && gi >= 0 && gi < g.nx - 1 && gj >= 0 && gj < g.ny - 1) {
      const int64_t u = a.ubase + ((int64_t)gi * (g.ny - 1) + gj) * g.nzseg + s;
      bv = a.voff[u] - vbase;
      bf = a.foff[u] - fbase;
      any = any || a.cnt[u] != 0;
    }
    sm.rbv[0][w][r] = bv;
    sm.rbf[0][w][r] = bf;
  }
  // a tile whose rows the count found empty writes nothing: skip it
  // before fetching any of its bricks (sparse fields have many)
  if (!__syncthreads_or(any)) return;

  // DMA: planes zc-1 .. zc+ZC, rows y0 .. y0+2, 64 samples from x0.  Per-lane
  // byte offsets relative to the brick origin are fixed; near the volume's
  // low faces the origin is outside the field, so the buffer starts at the
  // first sample that exists and lanes before it read 0 (out of range).
  constexpr int NROWS = NP * SY;
  constexpr int RPI = VEC ? 1024 / (4 * RD) : 1;
  constexpr int NI = VEC ? (NROWS + RPI - 1) / RPI : NROWS * DW;
  constexpr int NIW = (NI + kWaves - 1) / kWaves;
  uint32_t doff[NIW];
  bool dok[NIW];
#pragma unroll
  for (int m = 0; m < NIW; ++m) {
    const int ins = wv + m * kWaves;
    int row, byte_in_row;
    if (VEC) {
      row = ins * RPI + lane / (kWave / RPI);
      byte_in_row = (lane % (kWave / RPI)) * 16;
    } else {
      row = ins / DW;
      byte_in_row = (ins % DW) * 256 + lane * 4;
    }
    dok[m] = ins < NI && row < NROWS;
    const int pl = row / SY, y = row - pl * SY;
    doff[m] = (uint32_t)(((int64_t)pl * pstride + (int64_t)y * g.nx) * (int64_t)sizeof(T)) +
              (uint32_t)byte_in_row;
  }
  auto fetch = [&](int zc, int buf) {
    const int64_t start = (int64_t)(zc - 1) * pstride + (int64_t)y0 * g.nx + x0;
    const int64_t first = start < 0 ? 0 : start;
    const uint32_t skip = (uint32_t)((first - start) * (int64_t)sizeof(T));
    const int64_t left = (nelem - first) * (int64_t)sizeof(T);
    const __amdgpu_buffer_rsrc_t rs =
        buffer_rsrc(f + first, (uint32_t)(left < 0x7FFFFFFF ? left : 0x7FFFFFFF));
#pragma unroll
    for (int m = 0; m < NIW; ++m) {
      const int ins = wv + m * kWaves;
      // 0x80000000 is past every buffer's range: those lanes read zeros
      const uint32_t off = doff[m] >= skip ? doff[m] - skip : 0x80000000u;
      if (dok[m]) dma_to_lds<VEC ? 16 : 4>(rs, sm.smp[buf] + ins * (VEC ? 256 : kWave), off);
    }
    constexpr int NZ = (2 * NP + kWave - 1) / kWave;
    uint32_t* zd = (uint32_t*)sm.cz[buf];
    for (int ins = wv; ins < NZ; ins += kWaves) {
      const int p = ins * kWave + lane;
      if (p < 2 * NP) {
        const double* src = zp + min(max(zc - 1 + (p >> 1), 0), g.nz - 1);
        __builtin_amdgcn_global_load_lds((const void*)((const char*)src + 4 * (p & 1)),
                                         (void*)(zd + ins * kWave), 4, 0, 0);
      }
    }
  };

  uint64_t* vlist = sm.vlist[wv];
  uint64_t* flist = sm.flist[wv];
  uint16_t* lpv = sm.lpv[wv];
  uint16_t* lpf = sm.lpf[wv];
  uint64_t* vbits = sm.vbits[wv];
  uint64_t* fbits = sm.fbits[wv];
  for (int q = lane; q < VW; q += kWave) vbits[q] = 0;
  for (int q = lane; q < FW; q += kWave) fbits[q] = 0;
  const int h = lane / ZC, z = lane % ZC;
  const int wx0 = ib + wv * kTetWX - 1;  // global x of the wave's column xr = 0

  if (zs < ze) fetch(zs, 0);
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
  int buf = 0;
  for (int zc = zs; zc < ze; zc += ZC, buf ^= 1) {
    if (zc + ZC < ze) fetch(zc + ZC, buf ^ 1);
    const uint32_t* smp = sm.smp[buf];
    const double* czs = sm.cz[buf];
    const int64_t* rbv = sm.rbv[buf][wv];
    const int64_t* rbf = sm.rbf[buf][wv];
    int64_t* rbv_next = sm.rbv[buf ^ 1][wv];
    int64_t* rbf_next = sm.rbf[buf ^ 1][wv];

    for (int row = wv; row < NP * SY; row += kWaves) {
      const T v = *(const T*)(smp + row * RD + lane * DW);
      const uint64_t b = __ballot(v < thr);
      if (lane == 0) sm.mask[row] = b;
    }
    lds_barrier();

    int nvl = 0, nfl = 0;
    uint32_t V = 0, F = 0;
    auto flush = [&]() {
      int before = 0;
      for (uint32_t o0 = 0; o0 < V; o0 += kWave) {
        const uint32_t o = o0 + lane;
        const uint64_t word = vbits[o0 >> 6];
        const uint64_t upto = word & (~0ull >> (kWave - 1 - lane));
        const int k = before + __popcll(upto) - 1;
        before += __popcll(word);
        if (o >= V) break;
        const uint64_t e = vlist[k];
        const int xr = (int)(e & 15), c = (int)((e >> 4) & 15);
        const uint32_t cs = (uint32_t)((e >> 8) & 0xFF), cls = (uint32_t)((e >> 16) & 7);
        const uint32_t pv = (uint32_t)((e >> 19) & 0x1FF);
        const uint32_t q = o - lpv[k];
        const int col = wv * kTetWX + xr + 3;  // brick column of the cell
        const uint32_t* cell = smp + (c + 1) * (SY * RD) + 1 * RD + col * DW;
        double px, py, pz;
        if constexpr (METHOD == MT) {
          // class-0 voxels (all but the first layers) read the LDS copy
          uint32_t ed = sm.own0[cs * 8 + (q & 7)];
          if (cls != 0) ed = MT_OWN_EDGES[(cls * 256 + nat_to_mt(cs)) * 20 + q];
          const uint32_t ew = sm.edge[ed];
          const T sv = *(const T*)(cell + (ew & 0x3FF)), tv = *(const T*)(cell + ((ew >> 10) & 0x3FF));
          // vertPos (src/marching_tetrahedra.jl:42-55)
          const M rr = (iso - (M)sv) / (M)(tv - sv);
          const A aw = jl_min<A>(jl_max<A>((A)rr, eps), ome);
          const A bw = (A)1 - aw;
          const A wx = (A)((ew >> 20) & 1) * bw + (A)((ew >> 23) & 1) * aw;
          const A wy = (A)((ew >> 21) & 1) * bw + (A)((ew >> 24) & 1) * aw;
          const A wz = (A)((ew >> 22) & 1) * bw + (A)((ew >> 25) & 1) * aw;
          px = mt_axis<A>(cx32, wx, sm.cx[col], sm.cx[col + 1]);
          py = mt_axis<A>(cy32, wy, sm.cy[1], sm.cy[2]);
          pz = mt_axis<A>(cz32, wz, czs[c + 1], czs[c + 2]);
        } else {
          // the mean of the crossing points of the cell's 12 edges, each
          // interpolated as vertex_interp does (src/marching_cubes.jl:100-104),
          // summed per axis in promote(M, axis type)
          const uint32_t csm = nat_to_mc(cs);
          float fx = 0.f, fy = 0.f, fz = 0.f;
          double dx = 0.0, dy = 0.0, dz = 0.0;
          const bool sx = sizeof(M) == 4 && cx32, sy = sizeof(M) == 4 && cy32,
                     sz = sizeof(M) == 4 && cz32;
          int ncross = 0;
          for (int ed = 0; ed < 12; ++ed) {
            const uint32_t ew = sm.edge[ed];
            if (((csm >> ((ew >> 26) & 7)) & 1) == ((csm >> (ew >> 29)) & 1)) continue;
            ++ncross;
            const T va = *(const T*)(cell + (ew & 0x3FF)), vb = *(const T*)(cell + ((ew >> 10) & 0x3FF));
            const M mu = (iso - (M)va) / (M)(vb - va);
            const double qx = lerp_axis<M>(cx32, mu, sm.cx[col + ((ew >> 20) & 1)], sm.cx[col + ((ew >> 23) & 1)]);
            const double qy = lerp_axis<M>(cy32, mu, sm.cy[1 + ((ew >> 21) & 1)], sm.cy[1 + ((ew >> 24) & 1)]);
            const double qz = lerp_axis<M>(cz32, mu, czs[c + 1 + ((ew >> 22) & 1)], czs[c + 1 + ((ew >> 25) & 1)]);
            if (sx) fx = fx + (float)qx; else dx = dx + qx;
            if (sy) fy = fy + (float)qy; else dy = dy + qy;
            if (sz) fz = fz + (float)qz; else dz = dz + qz;
          }
          px = sx ? (double)(fx / (float)ncross) : dx / (double)ncross;
          py = sy ? (double)(fy / (float)ncross) : dy / (double)ncross;
          pz = sz ? (double)(fz / (float)ncross) : dz / (double)ncross;
        }
        const int64_t vi = rbv[16 + xr] + pv + q;
        if (vi < totv) {
          Out* out = verts + vi * 3;
          out[0] = (Out)px;
          out[1] = (Out)py;
          out[2] = (Out)pz;
        }
      }
      before = 0;
      for (uint32_t o0 = 0; o0 < F; o0 += kWave) {
        const uint32_t o = o0 + lane;
        const uint64_t word = fbits[o0 >> 6];
        const uint64_t upto = word & (~0ull >> (kWave - 1 - lane));
        const int k = before + __popcll(upto) - 1;
        before += __popcll(word);
        if (o >= F) break;
        const uint64_t e = flist[k];
        const int xr = (int)(e & 15), c = (int)((e >> 4) & 15);
        const uint32_t cs = (uint32_t)((e >> 8) & 0xFF);
        const uint32_t pf = (uint32_t)((e >> 28) & 0x1FF), pv = (uint32_t)((e >> 19) & 0x1FF);
        const uint32_t q = o - lpf[k];
        const int gi = wx0 + xr, gj = jb, gk = zc + c;
        const int r = 16 + xr;
        int64_t id[3];
        if constexpr (METHOD == MT) {
          const uint32_t fw = sm.face[cs * 12 + q];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            const uint32_t ed = (fw >> (5 * m)) & 31;
            const uint32_t ew = sm.edge[ed];
            // the owner is one cell back along each axis on which the edge's
            // end corner has coordinate 0 (when that cell exists)
            const int dx = (((ew >> 23) & 1) == 0) & (gi > 0);
            const int dy = (((ew >> 24) & 1) == 0) & (gj > 0);
            const int dz = (((ew >> 25) & 1) == 0) & (gk > 0);
            const uint32_t e2 = sm.shift[ed * 8 + (dx | (dy << 1) | (dz << 2))];
            const int r2 = r - dx - 16 * dy, z2 = c - dz;
            const uint32_t cls2 =
                (uint32_t)((gi - dx == 0) | ((gj - dy == 0) << 1) | ((gk - dz == 0) << 2));
            uint32_t oc;
            int64_t off;
            if (z2 >= 0) {
              oc = sm.ccase[wv][r2][z2];
              off = rbv[r2] + sm.cpv[wv][r2][z2];
            } else {
              oc = sm.pcase[wv][r2];
              off = rbv[r2] - (int64_t)sm.nown[cls2 * 256 + oc];
            }
            uint32_t rk = sm.rank0[oc * 19 + e2];
            if (cls2 != 0) rk = MT_RANK[(cls2 * 256 + nat_to_mt(oc)) * 19 + e2];
            id[m] = off + (int64_t)rk + 1;
          }
        } else {
          // quads around the crossing edges at corner 0, two triangles each,
          // axes in x, y, z order; the q-th triangle is triangle q&1 of the
          // (q>>1)-th quad
          const uint32_t csm = nat_to_mc(cs), c0 = csm & 1;
          const uint32_t axes = (uint32_t)((c0 != ((csm >> 1) & 1)) && gj >= 1 && gk >= 1) |
                                ((uint32_t)((c0 != ((csm >> 3) & 1)) && gk >= 1 && gi >= 1) << 1) |
                                ((uint32_t)((c0 != ((csm >> 4) & 1)) && gi >= 1 && gj >= 1) << 2);
          const uint32_t first = axes & (0u - axes);
          const uint32_t pick = (q >> 1) == 0 ? first
                              : (q >> 1) == 1 ? ((axes ^ first) & (0u - (axes ^ first))) : 4u;
          const int ax = pick == 1 ? 0 : pick == 2 ? 1 : 2;
          const int u = (ax + 1) % 3, w = (ax + 2) % 3;
          auto cell_id = [&](int du, int dw) -> int64_t {
            const int d0 = (u == 0 ? du : 0) + (w == 0 ? dw : 0);
            const int d1 = (u == 1 ? du : 0) + (w == 1 ? dw : 0);
            const int d2 = (u == 2 ? du : 0) + (w == 2 ? dw : 0);
            const int r2 = r - d0 - 16 * d1, z2 = c - d2;
            if (z2 >= 0) return rbv[r2] + sm.cpv[wv][r2][z2] + 1;
            return rbv[r2];  // the cell just below the chunk, active: its vertex is the row's last
          };
          const int64_t me = rbv[r] + pv + 1;
          const int64_t q1 = cell_id(1, 0), q2 = cell_id(1, 1), q3 = cell_id(0, 1);
          const int64_t a1 = c0 ? q1 : q3, a3 = c0 ? q3 : q1;
          id[0] = me;
          id[1] = (q & 1) ? q2 : a1;
          id[2] = (q & 1) ? a3 : q2;
        }
        const int64_t fi = rbf[r] + pf + q;
        if (fi < totf) {
          int64_t* out = faces + fi * 3;
          out[0] = id[0];
          out[1] = id[1];
          out[2] = id[2];
        }
      }
      wave_sync();
      for (uint32_t q = lane; q < (V + 63) / 64; q += kWave) vbits[q] = 0;
      for (uint32_t q = lane; q < (F + 63) / 64; q += kWave) fbits[q] = 0;
    };

    const int gk = zc + z;
    const bool zvalid = gk < ze;
    for (int st = 0; st < NSTEP; ++st) {
      const int r = st * R + h, xr = r & 15, yr = r >> 4;
      const int gi = wx0 + xr, gj = jb - 1 + yr;
      const bool exists = xr < 15 && gi >= 0 && gi < g.nx - 1 && gj >= 0 && gj < g.ny - 1;
      const bool emits = yr == 1 && xr >= 1;  // halo rows and columns only classify
      const int col = wv * kTetWX + xr + 3;
      const uint64_t* mz = sm.mask + (z + 1) * SY + yr;
      const uint32_t lo = (uint32_t)((mz[0] >> col) & 3) | ((uint32_t)((mz[1] >> col) & 3) << 2);
      const uint32_t hi =
          (uint32_t)((mz[SY] >> col) & 3) | ((uint32_t)((mz[SY + 1] >> col) & 3) << 2);
      const uint32_t cs = (exists && zvalid) ? (lo | (hi << 4)) : 0u;
      const uint32_t cls = (uint32_t)((gi == 0) | ((gj == 0) << 1) | ((gk == 0) << 2));
      uint32_t nv, nfc;
      if constexpr (METHOD == MT) {
        nv = sm.nown[cls * 256 + cs];
        nfc = emits ? (uint32_t)sm.nf[cs] : 0u;
      } else {
        const uint32_t csm = nat_to_mc(cs), c0 = csm & 1;
        nv = (cs != 0 && cs != 255) ? 1u : 0u;
        const uint32_t nq = (uint32_t)((c0 != ((csm >> 1) & 1)) && gj >= 1 && gk >= 1) +
                            (uint32_t)((c0 != ((csm >> 3) & 1)) && gk >= 1 && gi >= 1) +
                            (uint32_t)((c0 != ((csm >> 4) & 1)) && gi >= 1 && gj >= 1);
        nfc = (emits && nv) ? 2u * nq : 0u;
      }
      const uint32_t packed = nv | (nfc << 16);
      const uint32_t inc = seg_scan_dpp<ZC>(packed);
      const uint32_t ex_row = inc - packed;
      if constexpr (METHOD == MT) sm.ccase[wv][r][z] = (uint8_t)cs;
      sm.cpv[wv][r][z] = (uint16_t)(ex_row & 0xFFFF);
      uint32_t tots[R];
#pragma unroll
      for (int q = 0; q < R; ++q)
        tots[q] = (uint32_t)__builtin_amdgcn_readlane((int)inc, q * ZC + ZC - 1);
      uint32_t before = 0, mine = tots[0], sum = 0, vsum = 0;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int rq = st * R + q;
        const bool eq = ((rq >> 4) == 1) && ((rq & 15) >= 1);  // row q emits
        if (q == h) mine = tots[q];
        if (eq) {
          if (q < h) before += tots[q];
          sum += tots[q];
        }
        vsum += tots[q] & 0xFFFF;
      }
      if (z == 0) {
        if constexpr (METHOD == MT) {
          // the cell just below the chunk: its case, from planes 0 and 1
          const uint64_t* mb = sm.mask + yr;
          const uint32_t blo =
              (uint32_t)((mb[0] >> col) & 3) | ((uint32_t)((mb[1] >> col) & 3) << 2);
          const uint32_t bhi =
              (uint32_t)((mb[SY] >> col) & 3) | ((uint32_t)((mb[SY + 1] >> col) & 3) << 2);
          sm.pcase[wv][r] = (uint8_t)((exists && zc > 0) ? (blo | (bhi << 4)) : 0u);
        }
        rbv_next[r] = rbv[r] + (mine & 0xFFFF);
        rbf_next[r] = rbf[r] + (mine >> 16);
      }
      if (vsum == 0 && sum == 0) continue;
      const bool hasv = emits && (packed & 0xFFFF) != 0;
      const bool hasf = emits && (packed >> 16) != 0;
      const uint64_t balv = __ballot(hasv), balf = __ballot(hasf);
      const int nav = __popcll(balv), naf = __popcll(balf);
      if (nav + naf == 0) continue;
      if (nvl + nav > CAP || nfl + naf > CAP) {
        wave_sync();
        flush();
        wave_sync();
        nvl = 0;
        nfl = 0;
        V = 0;
        F = 0;
      }
      // slots of this step's emitting cells: (row, z) order = emission order
      const uint32_t ex = ex_row + before;
      const uint64_t ent = (uint64_t)xr | ((uint64_t)z << 4) | ((uint64_t)cs << 8) |
                           ((uint64_t)cls << 16) | ((uint64_t)(ex_row & 0xFFFF) << 19) |
                           ((uint64_t)(ex_row >> 16) << 28);
      if (hasv) {
        const int k = nvl + (int)lanes_below(balv);
        const uint32_t sv = V + (ex & 0xFFFF);
        vlist[k] = ent;
        lpv[k] = (uint16_t)sv;
        atomicOr((unsigned long long*)&vbits[sv >> 6], 1ull << (sv & 63));
      }
      if (hasf) {
        const int k = nfl + (int)lanes_below(balf);
        const uint32_t sf = F + (ex >> 16);
        flist[k] = ent;
        lpf[k] = (uint16_t)sf;
        atomicOr((unsigned long long*)&fbits[sf >> 6], 1ull << (sf & 63));
      }
      nvl += nav;
      nfl += naf;
      V += sum & 0xFFFF;
      F += sum >> 16;
    }
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
    flush();
    lds_barrier();
  }
}

template <int METHOD, typename T, typename M, typename A, typename Out, bool VEC, int CXM>
__global__ __launch_bounds__(kBlock) void k_emit_halo(VolArgs a) {
  emit_halo_tile<METHOD, T, M, A, Out, VEC, CXM>(a, xcd_tile(blockIdx.x, a.ntiles));
}

template <int METHOD, typename T, typename M, typename A, typename Out, bool VEC, int CXM>
__global__ __launch_bounds__(kBlock) void k_emit_halo_b(const VolArgs* __restrict__ vols,
                                                        const int32_t* __restrict__ emap,
                                                        int64_t gbase) {
  const int64_t b = gbase + blockIdx.x;
  const VolArgs a = vols[emap[b]];
  emit_halo_tile<METHOD, T, M, A, Out, VEC, CXM>(a, b - a.eblk0);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

int halo_variant(const Plan& p, VolArgs* a) {
  const int64_t esz = p.T == K_F32 ? 4 : 8;
  // brick DMA offsets are 32-bit and span ZC+2 planes
  if ((int64_t)(kTetZC + 2) * a->g.nx * a->g.ny * esz >= (1ll << 31)) return -1;
  a->nti = (a->g.nx - 1 + kTetTX - 1) / kTetTX;
  a->ntj = a->g.ny - 1;
  a->ntiles = (int64_t)a->nti * a->ntj * a->g.nzseg;
  // 16-byte DMA: rows 16-byte aligned (the brick starts 4 cells before a
  // tile whose x is a multiple of 56, so that start is aligned too)
  const bool vec = ((uintptr_t)a->f % 16 == 0) && (((int64_t)a->g.nx * esz) % 16 == 0);
  const int cxclass = a->cxmask == 0 ? 1 : a->cxmask == 7 ? 2 : 0;
  return (vec ? 0 : 1) | (cxclass << 1);
}

template <int METHOD, typename T, typename M, typename A, typename Out, int CXM>
static hipError_t halo_z(bool vec, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                         int64_t gbase, int64_t nblocks, hipStream_t st) {
  if (vols) {
    if (vec)
      hipLaunchKernelGGL((k_emit_halo_b<METHOD, T, M, A, Out, true, CXM>), dim3((unsigned)nblocks),
                         dim3(kBlock), 0, st, vols, emap, gbase);
    else
      hipLaunchKernelGGL((k_emit_halo_b<METHOD, T, M, A, Out, false, CXM>), dim3((unsigned)nblocks),
                         dim3(kBlock), 0, st, vols, emap, gbase);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_emit_halo<METHOD, T, M, A, Out, true, CXM>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_emit_halo<METHOD, T, M, A, Out, false, CXM>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

template <int METHOD, typename T, typename M, typename A, typename Out>
static hipError_t halo_t(int variant, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                         int64_t gbase, int64_t nblocks, hipStream_t st) {
  const bool vec = (variant & 1) == 0;
  switch ((variant >> 1) & 3) {
    case 1: return halo_z<METHOD, T, M, A, Out, 0>(vec, a, vols, emap, gbase, nblocks, st);
    case 2: return halo_z<METHOD, T, M, A, Out, 7>(vec, a, vols, emap, gbase, nblocks, st);
    default: return halo_z<METHOD, T, M, A, Out, -1>(vec, a, vols, emap, gbase, nblocks, st);
  }
}

static hipError_t halo_any(const Plan& p, int variant, const VolArgs& a, const VolArgs* vols,
                           const int32_t* emap, int64_t gbase, int64_t nblocks, hipStream_t st) {
#define MSH_ARGS variant, a, vols, emap, gbase, nblocks, st
  if (p.method == SN) {
    // no eps: A is M
    if (p.T == K_F32 && p.M == K_F32 && p.Out == K_F32)
      return halo_t<SN, float, float, float, float>(MSH_ARGS);
    if (p.T == K_F32 && p.M == K_F32) return halo_t<SN, float, float, float, double>(MSH_ARGS);
    if (p.T == K_F32) return halo_t<SN, float, double, double, double>(MSH_ARGS);
    return halo_t<SN, double, double, double, double>(MSH_ARGS);
  }
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32 && p.Out == K_F32)
    return halo_t<MT, float, float, float, float>(MSH_ARGS);
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32) return halo_t<MT, float, float, float, double>(MSH_ARGS);
  if (p.T == K_F32 && p.M == K_F32) return halo_t<MT, float, float, double, double>(MSH_ARGS);
  if (p.T == K_F32) return halo_t<MT, float, double, double, double>(MSH_ARGS);
  return halo_t<MT, double, double, double, double>(MSH_ARGS);
#undef MSH_ARGS
}

hipError_t launch_emit_halo(const Plan& p, int variant, const VolArgs& a, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
  return halo_any(p, variant, a, nullptr, nullptr, 0, 0, st);
}

hipError_t launch_emit_halo_batch(const Plan& p, int variant, const VolArgs* vols,
                                  const int32_t* emap, int64_t gbase, int64_t nblocks,
                                  hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  return halo_any(p, variant, VolArgs{}, vols, emap, gbase, nblocks, st);
}

}  // namespace msh
