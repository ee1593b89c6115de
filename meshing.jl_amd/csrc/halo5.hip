// halo5.hip -- Marching Tetrahedra (src/marching_tetrahedra.jl:105-163) and
// Naive Surface Nets (this library's definition, oracle/cpu_ref.cpp
// sn_isosurface) emit kernels with lanes along x.
//
// The layout of the Marching Cubes kernel with lanes along x (mc5.hip): lane
// i of wave w holds the cell row (x0 + i, jb + w) and marches it through the
// z chunk with running in-row prefixes.  A face of these methods names
// vertices owned by the cells at -x, -y and -z, so the block classifies one
// more row and column than it emits: wave 0 also classifies the row at -y,
// lane 3 holds the column at -x.  Every cell's case and in-row vertex prefix
// go to LDS (row slot 0 is the row at -y, slot w + 1 wave w's row), a
// barrier makes them visible, and the waves look the owners up there; the
// cell just below the chunk comes from one extra plane of the brick.
#include <stdlib.h>

#include <type_traits>

#include "device.h"

#define MSH_LUT static __constant__
#include "luts.h"

namespace msh {
namespace {

constexpr int kH5TX = 56;  // cells a block emits along x: lanes 4..59
constexpr int kH5CY = 4;   // cell rows a block emits along y: one per wave

// The tetrahedra emit's lookup tables in the layout its blocks keep in LDS,
// by natural case.  Built once per device (tables_setup in capi.cpp, at
// context creation) into a file-scope device copy; every block then copies the
// 15.5 KB image with 16-byte loads instead of gathering and repacking the
// constant tables itself (the gather took 12 of 162 us of a 256^3 gyroid
// emit).
struct MtTables {
  uint8_t nown[8 * 256];    // vertices created, by class and (natural) case
  uint8_t nf[256];          // faces, by case
  uint16_t face[256 * 12];  // edges of each face, 5 bits each
  uint8_t rank0[256 * 19];  // rank of each edge among a class-0 voxel's vertices
  uint8_t own0[256 * 8];    // edges of a class-0 voxel's vertices, in order
  // owner step of a face vertex, by edge and which of the cells at -x,
  // -y, -z exist (bits 0-2): the edge's index in its owner | the step
  // (dx, dy, dz) << 5
  uint8_t own_step[152];
  uint8_t pad[8];
};
static_assert(sizeof(MtTables) % 16 == 0, "copied in 16-byte words");
struct NoTables {
  uint8_t pad[16];
};

static __device__ __align__(16) MtTables g_mt_tables;

__global__ __launch_bounds__(kBlock) void k_mt_tables() {
  MtTables& t = g_mt_tables;
  for (int c = threadIdx.x; c < 256; c += kBlock) {
    const uint32_t m = nat_to_mt((uint32_t)c);
#pragma unroll
    for (int cls = 0; cls < 8; ++cls) t.nown[cls * 256 + c] = MT_NOWN[cls * 256 + m];
    const int nf = MT_NF[m];
    t.nf[c] = (uint8_t)nf;
    for (int q = 0; q < 12; ++q) {
      uint16_t w = 0;
      if (q < nf)
        w = (uint16_t)(MT_FACE[m * 36 + 3 * q] | (MT_FACE[m * 36 + 3 * q + 1] << 5) |
                       (MT_FACE[m * 36 + 3 * q + 2] << 10));
      t.face[c * 12 + q] = w;
    }
    for (int e = 0; e < 19; ++e) t.rank0[c * 19 + e] = MT_RANK[m * 19 + e];
    for (int q = 0; q < 8; ++q) t.own0[c * 8 + q] = q < 7 ? MT_OWN_EDGES[m * 20 + q] : 0;
  }
  for (int q = threadIdx.x; q < 152; q += kBlock) {
    // the owner is one cell back along each axis on which the edge's end
    // corner has coordinate 0, when that cell exists (bits of q & 7)
    const int ed = q >> 3;
    const int cb = MT_EDGE[ed] >> 4;
    const uint32_t d = ((mt_dx(cb) == 0 ? 1u : 0u) | (mt_dy(cb) == 0 ? 2u : 0u) |
                        (mt_dz(cb) == 0 ? 4u : 0u)) & (uint32_t)(q & 7);
    t.own_step[q] = (uint8_t)(MT_SHIFT[ed * 8 + d] | (d << 5));
  }
  if (threadIdx.x < 8) t.pad[threadIdx.x] = 0;
}

template <int METHOD, typename T, typename M, typename A, typename Out, int ZC, bool VEC, int CXM,
          int NBUF, class Src>
__device__ __forceinline__ void emit_h5_run(Src src) {
  static_assert(NBUF == 1 || NBUF == 2, "single- or double-buffered bricks");
  static_assert(METHOD == MT || METHOD == SN, "tetrahedra or surface nets");
  static_assert(ZC <= 16, "z fits 4 bits of a queue entry");
  constexpr int SX = 64, SY = kH5CY + 2, NP = ZC + 2;  // planes zc-1 .. zc+ZC, rows jb-1 .. jb+4
  constexpr int NS = kH5CY + 1;  // row slots: the row at -y, then one per wave
  constexpr int DW = (int)(sizeof(T) / 4), RD = SX * DW;
  constexpr int CAPE = 128, CAPV = 1024, CAPF = 1024;  // per wave
  static_assert(kH5TX * 13 <= CAPV && kH5TX * 12 <= CAPF && kH5TX <= CAPE, "one z step must fit");
  constexpr int K = METHOD == MT ? 1 : 0;
  struct Smem {
    uint32_t smp[NBUF][NP * SY * RD];
    double cz[NBUF][NP];
    double cx[SX], cy[SY];
    // tetrahedra only
    alignas(16) std::conditional_t<K != 0, MtTables, NoTables> mt;
    // in-row vertex prefix << 8 | case of each cell of the chunk (one LDS
    // word per owner look-up)
    uint32_t cpc[NS][K ? ZC : 1][64];
    uint8_t pc[NS][K ? 64 : 1];       // case of the cell just below the chunk
    uint32_t edge[19];                // per edge: corner offsets and axis bits
    uint16_t xmask[K ? 1 : 256];      // surface nets: crossing cube edges, by natural case
    uint16_t cp[NS][K ? 1 : ZC][64];  // surface nets: in-row vertex prefix of each cell
    int64_t rb[NS][64];               // row vertex offset at the chunk start
    int64_t dv[kWaves][64], df[kWaves][64];  // per row: output index minus queue slot
    // queue entries: x | z << 6 | case << 10 | class << 18 | first slot << 21
    uint32_t vlist[kWaves][CAPE], flist[kWaves][CAPE];
    uint64_t vbits[kWaves][CAPV / 64], fbits[kWaves][CAPF / 64];
    uint64_t tzm[2 * kWaves];  // the tile's z-block mask and write flag, per wave
  };
  __shared__ Smem sm;
  const int lane = lane_id(), wv = wave_id();
  // lookup tables: the same for every tile
  if constexpr (METHOD == MT) {
    const uint4* __restrict__ src = (const uint4*)&g_mt_tables;
    uint4* dst = (uint4*)&sm.mt;
    for (int q = threadIdx.x; q < (int)(sizeof(MtTables) / 16); q += kBlock) dst[q] = src[q];
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
  if constexpr (METHOD == SN) {
    for (int c = threadIdx.x; c < 256; c += kBlock) {
      const uint32_t csm = nat_to_mc((uint32_t)c);
      uint32_t m = 0;
      for (int ed = 0; ed < 12; ++ed) {
        const uint32_t ew = MC_EDGE[ed];
        m |= (((csm >> (ew & 0xF)) ^ (csm >> (ew >> 4))) & 1u) << ed;
      }
      sm.xmask[c] = (uint16_t)m;
    }
  }

  auto tile = [&](const VolArgs& a, const int64_t t) {
    const T* __restrict__ f = (const T*)a.f;
    Out* __restrict__ verts = (Out*)a.verts;
    int64_t* __restrict__ faces = a.faces;
    const Geom g = a.g;
    const int ib = (int)(t % a.nti) * kH5TX;
    const int64_t rest = t / a.nti;
    const int jb = (int)(rest % a.ntj) * kH5CY;
    const int s = (int)(rest / a.ntj);
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
    const int x0 = ib - 4, y0 = jb - 1;  // brick origin: sample column 0, row 0
    for (int q = threadIdx.x; q < SX; q += kBlock) sm.cx[q] = a.coords[min(max(x0 + q, 0), g.nx - 1)];
    if (threadIdx.x < SY)
      sm.cy[threadIdx.x] = a.coords[g.nx + min(max(y0 + (int)threadIdx.x, 0), g.ny - 1)];

    // this lane's cell row (row slot sl), and for wave 0 the row at -y
    const int sl = wv + 1;
    const int gi = x0 + lane, gj = y0 + sl;
    const bool colok = lane < SX - 1 && gi >= 0 && gi < g.nx - 1;
    const bool exists = colok && gj < g.ny - 1;
    const bool emits = exists && lane >= 4 && lane < 4 + kH5TX;
    const bool hexists = wv == 0 && colok && y0 >= 0;
    int64_t rbv = 0, rbf = 0, hrbv = 0;
    bool writes = false;
    uint64_t zm = 0;
    if (exists) {
      const int64_t u = a.ubase + ((int64_t)gi * (g.ny - 1) + gj) * g.nzseg + s;
      const uint64_t zu = zmask_at<T>(a, gi, gj, s);  // loaded beside the offsets
      const uint32_t cu = a.cnt[u];
      rbv = a.voff[u] - vbase;
      rbf = a.foff[u] - fbase;
      writes = emits && cu != 0;
      zm = cu != 0 ? zu : 0ull;
    }
    if (hexists) {
      const int64_t u = a.ubase + ((int64_t)gi * (g.ny - 1) + y0) * g.nzseg + s;
      const uint64_t zu = zmask_at<T>(a, gi, y0, s);
      hrbv = a.voff[u] - vbase;
      zm |= a.cnt[u] != 0 ? zu : 0ull;
    }
    // a tile whose emitting rows the count found empty writes nothing (the
    // rows around it only supply owners to faces): skip it before fetching
    // any of its bricks.  Otherwise only the z blocks where some row of the
    // brick, the rows around it included, may hold an active cell are
    // walked: the chunks of the other blocks change no row's vertex offset,
    // write nothing and are never fetched.
    bool any;
    const uint64_t tmask = tile_zmask(zm, writes, sm.tzm, &any);
    if (!any || tmask == 0) return;  // (a writing row marks a block)
    const int zsh = zmask_shift(g.zseg);
    const int z0 = zs + (__builtin_ctzll(tmask) << zsh);

    // DMA: planes zc-1 .. zc+ZC, rows y0 .. y0+4, 64 samples from x0.  Near
    // the volume's low faces the origin is outside the field: the buffer starts
    // at the first sample that exists and lanes before it read 0.
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

    uint32_t* vlist = sm.vlist[wv];
    uint32_t* flist = sm.flist[wv];
    uint64_t* vbits = sm.vbits[wv];
    uint64_t* fbits = sm.fbits[wv];
    for (int q = lane; q < CAPV / 64; q += kWave) vbits[q] = 0;
    for (int q = lane; q < CAPF / 64; q += kWave) fbits[q] = 0;

    if (z0 < ze) fetch(z0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    lds_barrier();
    int buf = 0;
    for (int zc = z0, znx; zc < ze; zc = znx, buf ^= NBUF - 1) {
      znx = next_zchunk(zc, ZC, zs, ze, zsh, tmask);
      // double-buffered: the next brick streams in during this chunk;
      // single-buffered: it is fetched after the chunk (more blocks per CU)
      if (NBUF == 2 && znx < ze) fetch(znx, buf ^ 1);
      const uint32_t* smp = sm.smp[buf];
      const double* czs = sm.cz[buf];

      // classify this lane's row: natural-order face nibble per plane
      uint32_t nib[NP];
  #pragma unroll
      for (int p = 0; p < NP; ++p) {
        const T v0 = *(const T*)(smp + (p * SY + sl) * RD + lane * DW);
        const T v1 = *(const T*)(smp + (p * SY + sl + 1) * RD + lane * DW);
        const uint32_t w = (uint32_t)(v0 < thr) | ((uint32_t)(v1 < thr) << 2);
        nib[p] = w | (next_lane(w, 0u) << 1);
      }
      const int zn = min(ZC, ze - zc);
      uint32_t cs[ZC], nvz[ZC], nfz[ZC];
      uint32_t runv = 0, tot = 0;  // in-row vertex prefix; emitted vertices | faces << 16
  #pragma unroll
      for (int z = 0; z < ZC; ++z) {
        const int gk = zc + z;
        cs[z] = (exists && z < zn) ? (nib[z + 1] | (nib[z + 2] << 4)) : 0u;
        const uint32_t cls = (uint32_t)((gi == 0) | ((gj == 0) << 1) | ((gk == 0) << 2));
        if constexpr (METHOD == MT) {
          nvz[z] = sm.mt.nown[cls * 256 + cs[z]];
          nfz[z] = emits ? (uint32_t)sm.mt.nf[cs[z]] : 0u;
          sm.cpc[sl][z][lane] = (runv << 8) | cs[z];
        } else {
          const uint32_t csm = nat_to_mc(cs[z]), c0 = csm & 1;
          nvz[z] = (cs[z] != 0 && cs[z] != 255) ? 1u : 0u;
          const uint32_t nq = (uint32_t)((c0 != ((csm >> 1) & 1)) && gj >= 1 && gk >= 1) +
                              (uint32_t)((c0 != ((csm >> 3) & 1)) && gk >= 1 && gi >= 1) +
                              (uint32_t)((c0 != ((csm >> 4) & 1)) && gi >= 1 && gj >= 1);
          nfz[z] = (emits && nvz[z]) ? 2u * nq : 0u;
          sm.cp[sl][z][lane] = (uint16_t)runv;
        }
        runv += nvz[z];
        if (emits) tot += nvz[z] | (nfz[z] << 16);
      }
      if constexpr (METHOD == MT) sm.pc[sl][lane] = (uint8_t)((exists && zc > 0) ? (nib[0] | (nib[1] << 4)) : 0u);
      sm.rb[sl][lane] = rbv;
      if (wv == 0) {
        // the row at -y (slot 0): cases and vertex prefixes for the owner
        // lookups of row jb's faces; it emits nothing here
        uint32_t hn[NP];
  #pragma unroll
        for (int p = 0; p < NP; ++p) {
          const T v0 = *(const T*)(smp + (p * SY) * RD + lane * DW);
          const T v1 = *(const T*)(smp + (p * SY + 1) * RD + lane * DW);
          const uint32_t w = (uint32_t)(v0 < thr) | ((uint32_t)(v1 < thr) << 2);
          hn[p] = w | (next_lane(w, 0u) << 1);
        }
        uint32_t hrun = 0;
  #pragma unroll
        for (int z = 0; z < ZC; ++z) {
          const uint32_t c = (hexists && z < zn) ? (hn[z + 1] | (hn[z + 2] << 4)) : 0u;
          uint32_t nv;
          if constexpr (METHOD == MT) {
            const uint32_t cls = (uint32_t)((gi == 0) | ((y0 == 0) << 1) | ((zc + z == 0) << 2));
            nv = sm.mt.nown[cls * 256 + c];
            sm.cpc[0][z][lane] = (hrun << 8) | c;
          } else {
            nv = (c != 0 && c != 255) ? 1u : 0u;
            sm.cp[0][z][lane] = (uint16_t)hrun;
          }
          hrun += nv;
        }
        if constexpr (METHOD == MT) sm.pc[0][lane] = (uint8_t)((hexists && zc > 0) ? (hn[0] | (hn[1] << 4)) : 0u);
        sm.rb[0][lane] = hrbv;
        hrbv += hrun;
      }
      // this wave's share of the next brick must land before the barrier that
      // ends the chunk; wait now, before the queue's stores are issued
      __builtin_amdgcn_s_waitcnt(0);
      lds_barrier();  // every row's cases and prefixes are in LDS

      {
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
            const uint32_t e = vlist[k];
            const int lc = (int)(e & 63), c = (int)((e >> 6) & 15);
            const uint32_t ecs = (e >> 10) & 0xFF, cls = (e >> 18) & 7, sv = e >> 21;
            const uint32_t q = o - sv;
            const uint32_t* cell = smp + ((c + 1) * SY + sl) * RD + lc * DW;
            double px, py, pz;
            if constexpr (METHOD == MT) {
              // class-0 voxels (all but the first layers) read the LDS copy
              uint32_t ed = sm.mt.own0[ecs * 8 + (q & 7)];
              if (cls != 0) ed = MT_OWN_EDGES[(cls * 256 + nat_to_mt(ecs)) * 20 + q];
              const uint32_t ew = sm.edge[ed];
              const T sv0 = *(const T*)(cell + (ew & 0x3FF)), tv = *(const T*)(cell + ((ew >> 10) & 0x3FF));
              // vertPos (src/marching_tetrahedra.jl:42-55)
              const M rr = (iso - (M)sv0) / (M)(tv - sv0);
              const A aw = jl_min<A>(jl_max<A>((A)rr, eps), ome);
              const A bw = (A)1 - aw;
              const A wx = (A)((ew >> 20) & 1) * bw + (A)((ew >> 23) & 1) * aw;
              const A wy = (A)((ew >> 21) & 1) * bw + (A)((ew >> 24) & 1) * aw;
              const A wz = (A)((ew >> 22) & 1) * bw + (A)((ew >> 25) & 1) * aw;
              px = mt_axis<A>(cx32, wx, sm.cx[lc], sm.cx[lc + 1]);
              py = mt_axis<A>(cy32, wy, sm.cy[sl], sm.cy[sl + 1]);
              pz = mt_axis<A>(cz32, wz, czs[c + 1], czs[c + 2]);
            } else {
              // the mean of the crossing points of the cell's 12 edges, each
              // interpolated as vertex_interp does (src/marching_cubes.jl:100-104),
              // summed per axis in promote(M, axis type)
              float fx = 0.f, fy = 0.f, fz = 0.f;
              double dx = 0.0, dy = 0.0, dz = 0.0;
              const bool sx = sizeof(M) == 4 && cx32, sy = sizeof(M) == 4 && cy32,
                         sz = sizeof(M) == 4 && cz32;
              // the crossing edges in edge order (the order of the sum), one
              // iteration each rather than a pass over all 12
              uint32_t xm = sm.xmask[ecs];
              const int ncross = __popc(xm);
              while (xm != 0) {
                const int ed = __builtin_ctz(xm);
                xm &= xm - 1;
                const uint32_t ew = sm.edge[ed];
                const T va = *(const T*)(cell + (ew & 0x3FF)), vb = *(const T*)(cell + ((ew >> 10) & 0x3FF));
                const M mu = (iso - (M)va) / (M)(vb - va);
                const double qx = lerp_axis<M>(cx32, mu, sm.cx[lc + ((ew >> 20) & 1)], sm.cx[lc + ((ew >> 23) & 1)]);
                const double qy = lerp_axis<M>(cy32, mu, sm.cy[sl + ((ew >> 21) & 1)], sm.cy[sl + ((ew >> 24) & 1)]);
                const double qz = lerp_axis<M>(cz32, mu, czs[c + 1 + ((ew >> 22) & 1)], czs[c + 1 + ((ew >> 25) & 1)]);
                if (sx) fx = fx + (float)qx; else dx = dx + qx;
                if (sy) fy = fy + (float)qy; else dy = dy + qy;
                if (sz) fz = fz + (float)qz; else dz = dz + qz;
              }
              px = sx ? (double)(fx / (float)ncross) : dx / (double)ncross;
              py = sy ? (double)(fy / (float)ncross) : dy / (double)ncross;
              pz = sz ? (double)(fz / (float)ncross) : dz / (double)ncross;
            }
            const int64_t vi = sm.dv[wv][lc] + o;
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
            const uint32_t e = flist[k];
            const int lc = (int)(e & 63), c = (int)((e >> 6) & 15);
            const uint32_t ecs = (e >> 10) & 0xFF, sf = e >> 21;
            const uint32_t q = o - sf;
            const int ci = x0 + lc, cj = gj, ck = zc + c;
            int64_t id[3];
            if constexpr (METHOD == MT) {
              const uint32_t fw = sm.mt.face[ecs * 12 + q];
              const uint32_t nb = (uint32_t)((ci > 0) | ((cj > 0) << 1) | ((ck > 0) << 2));
  #pragma unroll
              for (int m = 0; m < 3; ++m) {
                const uint32_t ed = (fw >> (5 * m)) & 31;
                // the owner (src/marching_tetrahedra.jl:67-84: the first voxel
                // to request the edge) and the edge's index there
                const uint32_t st = sm.mt.own_step[ed * 8 + nb];
                const uint32_t e2 = st & 31;
                const int dx = (st >> 5) & 1, dy = (st >> 6) & 1, dz = (st >> 7) & 1;
                const int l2 = lc - dx, w2 = sl - dy, z2 = c - dz;
                const uint32_t cls2 =
                    (uint32_t)((ci - dx == 0) | ((cj - dy == 0) << 1) | ((ck - dz == 0) << 2));
                uint32_t oc;
                int64_t off;
                if (z2 >= 0) {
                  const uint32_t w = sm.cpc[w2][z2][l2];
                  oc = w & 0xFF;
                  off = sm.rb[w2][l2] + (int64_t)(w >> 8);
                } else {
                  oc = sm.pc[w2][l2];
                  off = sm.rb[w2][l2] - (int64_t)sm.mt.nown[cls2 * 256 + oc];
                }
                uint32_t rk = sm.mt.rank0[oc * 19 + e2];
                if (cls2 != 0) rk = MT_RANK[(cls2 * 256 + nat_to_mt(oc)) * 19 + e2];
                id[m] = off + (int64_t)rk + 1;
              }
            } else {
              // quads around the crossing edges at corner 0, two triangles each,
              // axes in x, y, z order; the q-th triangle is triangle q&1 of the
              // (q>>1)-th quad
              const uint32_t csm = nat_to_mc(ecs), c0 = csm & 1;
              const uint32_t axes = (uint32_t)((c0 != ((csm >> 1) & 1)) && cj >= 1 && ck >= 1) |
                                    ((uint32_t)((c0 != ((csm >> 3) & 1)) && ck >= 1 && ci >= 1) << 1) |
                                    ((uint32_t)((c0 != ((csm >> 4) & 1)) && ci >= 1 && cj >= 1) << 2);
              const uint32_t first = axes & (0u - axes);
              const uint32_t pick = (q >> 1) == 0 ? first
                                  : (q >> 1) == 1 ? ((axes ^ first) & (0u - (axes ^ first))) : 4u;
              const int ax = pick == 1 ? 0 : pick == 2 ? 1 : 2;
              const int u = (ax + 1) % 3, w = (ax + 2) % 3;
              auto cell_id = [&](int du, int dw) -> int64_t {
                const int d0 = (u == 0 ? du : 0) + (w == 0 ? dw : 0);
                const int d1 = (u == 1 ? du : 0) + (w == 1 ? dw : 0);
                const int d2 = (u == 2 ? du : 0) + (w == 2 ? dw : 0);
                const int l2 = lc - d0, w2 = sl - d1, z2 = c - d2;
                if (z2 >= 0) return sm.rb[w2][l2] + sm.cp[w2][z2][l2] + 1;
                return sm.rb[w2][l2];  // the cell just below the chunk, active: the row's last vertex
              };
              const int64_t me = sm.rb[sl][lc] + sm.cp[sl][c][lc] + 1;
              const int64_t q1 = cell_id(1, 0), q2 = cell_id(1, 1), q3 = cell_id(0, 1);
              const int64_t a1 = c0 ? q1 : q3, a3 = c0 ? q3 : q1;
              id[0] = me;
              id[1] = (q & 1) ? q2 : a1;
              id[2] = (q & 1) ? a3 : q2;
            }
            const int64_t fi = sm.df[wv][lc] + o;
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

        // queue the emitting cells z in [za, zb) of every row of this wave,
        // rows one after another, each row's cells in z order
        auto append = [&](int za, int zb, uint32_t rtot, uint32_t run0) {
          uint32_t nev = 0, nef = 0;
  #pragma unroll
          for (int z = 0; z < ZC; ++z) {
            const bool in = z >= za && z < zb && emits;
            nev += (in && nvz[z] != 0) ? 1u : 0u;
            nef += (in && nfz[z] != 0) ? 1u : 0u;
          }
          const uint32_t inct = wave_scan_dpp(rtot), incv = wave_scan_dpp(nev), incf = wave_scan_dpp(nef);
          const uint32_t sumt = (uint32_t)__builtin_amdgcn_readlane((int)inct, kWave - 1);
          const uint32_t sumv = (uint32_t)__builtin_amdgcn_readlane((int)incv, kWave - 1);
          const uint32_t sumf = (uint32_t)__builtin_amdgcn_readlane((int)incf, kWave - 1);
          const uint32_t bv = V + ((inct - rtot) & 0xFFFF), bf = F + ((inct - rtot) >> 16);
          uint32_t kv = (uint32_t)nvl + incv - nev, kf = (uint32_t)nfl + incf - nef;
          if (nev | nef) {
            sm.dv[wv][lane] = rbv + (int64_t)(run0 & 0xFFFF) - (int64_t)bv;
            sm.df[wv][lane] = rbf + (int64_t)(run0 >> 16) - (int64_t)bf;
          }
          uint32_t run = 0;
  #pragma unroll
          for (int z = 0; z < ZC; ++z) {
            const bool in = z >= za && z < zb && emits;
            if (in) {
              const uint32_t cls = (uint32_t)((gi == 0) | ((gj == 0) << 1) | ((zc + z == 0) << 2));
              const uint32_t head = (uint32_t)lane | ((uint32_t)z << 6) | (cs[z] << 10) | (cls << 18);
              if (nvz[z] != 0) {
                const uint32_t sv = bv + (run & 0xFFFF);
                vlist[kv++] = head | (sv << 21);
                atomicOr((unsigned long long*)&vbits[sv >> 6], 1ull << (sv & 63));
              }
              if (nfz[z] != 0) {
                const uint32_t sf = bf + (run >> 16);
                flist[kf++] = head | (sf << 21);
                atomicOr((unsigned long long*)&fbits[sf >> 6], 1ull << (sf & 63));
              }
              run += nvz[z] | (nfz[z] << 16);
            }
          }
          nvl += (int)sumv;
          nfl += (int)sumf;
          V += sumt & 0xFFFF;
          F += sumt >> 16;
        };

        const uint32_t ttot = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_dpp(tot), kWave - 1);
        uint32_t cv = 0, cf = 0;
  #pragma unroll
        for (int z = 0; z < ZC; ++z) {
          cv += (emits && nvz[z] != 0) ? 1u : 0u;
          cf += (emits && nfz[z] != 0) ? 1u : 0u;
        }
        const uint32_t tcv = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_dpp(cv), kWave - 1);
        const uint32_t tcf = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_dpp(cf), kWave - 1);
        // the whole chunk in one go when it fits the queue, else one z step at a time
        const bool whole = (ttot & 0xFFFF) <= (uint32_t)CAPV && (ttot >> 16) <= (uint32_t)CAPF &&
                           tcv <= (uint32_t)CAPE && tcf <= (uint32_t)CAPE;
        uint32_t run0 = 0;
        for (int za = 0; ttot != 0 && za < zn;) {
          const int zb = whole ? ZC : za + 1;
          uint32_t part = 0;
  #pragma unroll
          for (int q = 0; q < ZC; ++q)
            part += (emits && q >= za && q < zb) ? (nvz[q] | (nfz[q] << 16)) : 0u;
          append(za, zb, part, run0);
          run0 += part;
          wave_sync();
          flush();
          wave_sync();
          nvl = 0;
          nfl = 0;
          V = 0;
          F = 0;
          za = zb;
        }
      }
      rbv += runv;
      rbf += tot >> 16;
      lds_barrier();
      if (NBUF == 1 && znx < ze) {
        fetch(znx, 0);
        __builtin_amdgcn_s_waitcnt(0);
        lds_barrier();
      }
    }
  };
  if constexpr (Src::kOne) {
    tile(*src.p, src.t);
  } else {
    VolArgs a;
    int64_t t;
    while (src(a, t)) tile(a, t);
  }
}

template <int METHOD, typename T, typename M, typename A, typename Out, int ZC, bool VEC, int CXM,
          int NBUF>
__global__ __launch_bounds__(kBlock) void k_emit_h5(VolArgs a) {
  emit_h5_run<METHOD, T, M, A, Out, ZC, VEC, CXM, NBUF>(
      OneTile{&a, xcd_tile(blockIdx.x, a.ntiles), false});
}

// Batched, persistent: the launch's blocks take the tiles of its volumes
// from a work queue (TileQueue).
template <int METHOD, typename T, typename M, typename A, typename Out, int ZC, bool VEC, int CXM,
          int NBUF>
__global__ __launch_bounds__(kBlock) void k_emit_h5_b(const VolArgs* __restrict__ vols,
                                                      const int32_t* __restrict__ emap,
                                                      int64_t gbase, int64_t ntiles,
                                                      uint32_t* __restrict__ ctr) {
  emit_h5_run<METHOD, T, M, A, Out, ZC, VEC, CXM, NBUF>(
      TileQueue<true>{vols, emap, gbase, ntiles, ctr, 0});
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

hipError_t launch_mt_tables(hipStream_t st) {
  hipLaunchKernelGGL(k_mt_tables, dim3(1), dim3(kBlock), 0, st);
  return hipGetLastError();
}

// z chunk: 8 cells for surface nets (16 measured slower on dense fields:
// more chunks overflow the queue) and 4 for tetrahedra, whose LDS tables
// leave room for two blocks per CU at 8 but three at 4 (Float32, 512^3 value
// noise: 2.42 against 2.81 ms; 256^3 gyroid 0.178 against 0.182 ms).  The
// half-depth bricks of surface nets (family 3: chunks of 4, three blocks per
// CU become four) win on dense fields (512^3 value noise 1.29 against 1.43
// ms) and lose on sparse ones (256^3 gyroid 0.147 against 0.131 ms), so the
// autotuner chooses.
template <int METHOD, typename T, bool HALF = false>
static constexpr int h5_zc() { return METHOD == MT ? 4 : (HALF ? 4 : 8); }
// bricks per block: 2 for tetrahedra (the next chunk's brick streams in
// during this one), 1 for surface nets, whose lighter LDS then holds four
// blocks per CU instead of three (256^3 gyroid 0.115 against 0.130 ms,
// 512^3 value noise 1.26 against 1.30 ms; tetrahedra single-buffered were
// slower, 2.49 against 2.36 ms)
template <int METHOD, typename T, bool HALF = false>
static constexpr int h5_nbuf() { return METHOD == SN ? 1 : 2; }

int h5_variant(const Plan& p, VolArgs* a) {
  const bool half = p.family == 3 && p.method == SN;
  const int zc = p.method == SN ? (half ? h5_zc<SN, float, true>() : h5_zc<SN, float>())
                                : h5_zc<MT, float>();
  const int64_t esz = p.T == K_F32 ? 4 : 8;
  // the brick's 32-bit DMA offsets reach plane zc+1, row kH5CY+1, sample 63
  if (((int64_t)(zc + 1) * a->g.nx * a->g.ny + (int64_t)(kH5CY + 1) * a->g.nx + 64) * esz >=
      (1ll << 31))
    return -1;
  a->nti = (a->g.nx - 1 + kH5TX - 1) / kH5TX;
  a->ntj = (a->g.ny - 1 + kH5CY - 1) / kH5CY;
  a->ntiles = (a->g.nx > 1 && a->g.ny > 1) ? (int64_t)a->nti * a->ntj * a->g.nzseg : 0;
  // 16-byte DMA: rows 16-byte aligned (the brick starts 4 cells before a
  // tile whose x is a multiple of 56, so that start is aligned too)
  const bool vec = ((uintptr_t)a->f % 16 == 0) && (((int64_t)a->g.nx * esz) % 16 == 0);
  const int cxclass = a->cxmask == 0 ? 1 : a->cxmask == 7 ? 2 : 0;
  return (vec ? 0 : 1) | (cxclass << 1) | (half ? 8 : 0);
}

template <int METHOD, typename T, typename M, typename A, typename Out, int CXM, bool HALF = false>
static hipError_t h5_z(bool vec, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                       int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  constexpr int ZC = h5_zc<METHOD, T, HALF>();
  constexpr int NB = h5_nbuf<METHOD, T, HALF>();
  if (vols) {
    const auto k = vec ? k_emit_h5_b<METHOD, T, M, A, Out, ZC, true, CXM, NB>
                       : k_emit_h5_b<METHOD, T, M, A, Out, ZC, false, CXM, NB>;
    hipLaunchKernelGGL(k, dim3(persistent_grid((const void*)k, nblocks)), dim3(kBlock), 0, st,
                       vols, emap, gbase, nblocks, ctr);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_emit_h5<METHOD, T, M, A, Out, ZC, true, CXM, NB>),
                         dim3((unsigned)a.ntiles), dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_emit_h5<METHOD, T, M, A, Out, ZC, false, CXM, NB>),
                         dim3((unsigned)a.ntiles), dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

template <int METHOD, typename T, typename M, typename A, typename Out>
static hipError_t h5_t(int variant, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                       int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  const bool vec = (variant & 1) == 0;
  if constexpr (METHOD == SN) {
    if (variant & 8) {
      switch ((variant >> 1) & 3) {
        case 1: return h5_z<SN, T, M, A, Out, 0, true>(vec, a, vols, emap, gbase, nblocks, ctr, st);
        case 2: return h5_z<SN, T, M, A, Out, 7, true>(vec, a, vols, emap, gbase, nblocks, ctr, st);
        default:
          return h5_z<SN, T, M, A, Out, -1, true>(vec, a, vols, emap, gbase, nblocks, ctr, st);
      }
    }
  }
  switch ((variant >> 1) & 3) {
    case 1: return h5_z<METHOD, T, M, A, Out, 0>(vec, a, vols, emap, gbase, nblocks, ctr, st);
    case 2: return h5_z<METHOD, T, M, A, Out, 7>(vec, a, vols, emap, gbase, nblocks, ctr, st);
    default: return h5_z<METHOD, T, M, A, Out, -1>(vec, a, vols, emap, gbase, nblocks, ctr, st);
  }
}

static hipError_t h5_any(const Plan& p, int variant, const VolArgs& a, const VolArgs* vols,
                         const int32_t* emap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                         hipStream_t st) {
#define MSH_ARGS variant, a, vols, emap, gbase, nblocks, ctr, st
  if (p.method == SN) {
    // no eps: A is M
    if (p.T == K_F32 && p.M == K_F32 && p.Out == K_F32)
      return h5_t<SN, float, float, float, float>(MSH_ARGS);
    if (p.T == K_F32 && p.M == K_F32) return h5_t<SN, float, float, float, double>(MSH_ARGS);
    if (p.T == K_F32) return h5_t<SN, float, double, double, double>(MSH_ARGS);
    return h5_t<SN, double, double, double, double>(MSH_ARGS);
  }
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32 && p.Out == K_F32)
    return h5_t<MT, float, float, float, float>(MSH_ARGS);
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32) return h5_t<MT, float, float, float, double>(MSH_ARGS);
  if (p.T == K_F32 && p.M == K_F32) return h5_t<MT, float, float, double, double>(MSH_ARGS);
  if (p.T == K_F32) return h5_t<MT, float, double, double, double>(MSH_ARGS);
  return h5_t<MT, double, double, double, double>(MSH_ARGS);
#undef MSH_ARGS
}

hipError_t launch_emit_h5(const Plan& p, int variant, const VolArgs& a, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
  return h5_any(p, variant, a, nullptr, nullptr, 0, 0, nullptr, st);
}

hipError_t launch_emit_h5_batch(const Plan& p, int variant, const VolArgs* vols,
                                const int32_t* emap, int64_t gbase, int64_t nblocks,
                                uint32_t* ctr, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  return h5_any(p, variant, VolArgs{}, vols, emap, gbase, nblocks, ctr, st);
}

}  // namespace msh
