// mc5.hip -- the Marching Cubes emit kernel with lanes along x
// (src/marching_cubes.jl:27-92).
//
// A block owns 60 cells along x and 4 cell rows along y (one per wave) and
// walks its z segment in chunks of ZC cells.  Per chunk:
//  * the brick (ZC+1 planes x 5 sample rows x 64 samples) arrives in LDS by
//    DMA, double-buffered;
//  * lane i of wave w holds row (i, w): it classifies the row's ZC cells from
//    its own two samples per plane and its right neighbour's (one DPP move),
//    looks up their vertex and face counts and keeps the in-row prefix as a
//    running sum -- the reference's emission order inside a row is z order;
//  * one wave scan of the rows' chunk totals places every row's outputs
//    contiguously in the wave's queue, and the list of active cells is
//    emitted slot-parallel (lane o writes queue slot o), so consecutive lanes
//    write consecutive outputs of a row.
// A row's output offset lives in the lane's registers from chunk to chunk.
// A chunk whose outputs do not fit the queue is taken one z step at a time.
#include <stdlib.h>

#include <type_traits>

#include "device.h"

#define MSH_LUT static __constant__
#include "luts.h"

namespace msh {
namespace {

constexpr int kMc5TX = 60;  // cells along x per block (lanes 60..63 only load)
constexpr int kMc5CY = 4;   // cell rows along y per block: one per wave

template <typename T, typename M, typename Out, int ZC, bool VEC, int CXM, int NBUF, class Src>
__device__ __forceinline__ void emit_mc5_run(Src src) {
  static_assert(NBUF == 1 || NBUF == 2, "single- or double-buffered bricks");
  constexpr int TX = kMc5TX, SX = 64, SY = kMc5CY + 1;
  constexpr int NP = ZC + 1;
  constexpr int DW = (int)(sizeof(T) / 4);
  constexpr int RD = SX * DW;  // dwords per sample row
  constexpr int CAPE = 256, CAPV = 1024, CAPF = 512;  // per wave: entries, vertex and face slots
  static_assert(TX * 12 <= CAPV && TX * 5 <= CAPF && TX <= CAPE, "one z step must fit");
  struct Smem {
    uint32_t smp[NBUF][NP * SY * RD];
    double cz[NBUF][NP];
    double cx[SX], cy[SY];
    uint32_t cnt[256];
    uint64_t lv[256], lt[256];
    uint32_t le[16];                     // per edge: brick offsets of its corners, (dx,dy,dz)
    int64_t dv[kWaves][64], df[kWaves][64];  // per row: output index minus queue slot
    uint32_t list[kWaves][CAPE];         // x | z << 6 | case << 10 | first vertex slot << 18
    uint16_t lpf[kWaves][CAPE];          // first face slot
    uint64_t vbits[kWaves][CAPV / 64], fbits[kWaves][CAPF / 64];  // slot o starts an entry
    uint64_t tzm[2 * kWaves];            // the tile's z-block mask and write flag, per wave
  };
  __shared__ Smem sm;
  const int lane = lane_id(), wv = wave_id();
  // lookup tables: the same for every tile
  for (int c = threadIdx.x; c < 256; c += kBlock) {
    const uint32_t m = nat_to_mc((uint32_t)c);
    const uint64_t vw = MC_CASE_VERTS[m], tw = MC_CASE_TRIS[m];
    sm.cnt[c] = (uint32_t)(vw & 0xF) | ((uint32_t)(tw & 0x7) << 16);
    sm.lv[c] = vw;
    sm.lt[c] = tw;
  }
  if (threadIdx.x < 12) {
    // edge e runs from corner a to corner b (src/lut/mc.jl _mc_edge_list);
    // word: offset of a (dwords from the cell's corner 0 in the brick) |
    // offset of b << 10 | dx,dy,dz of a << 20 | dx,dy,dz of b << 23
    const uint32_t ew = MC_EDGE[threadIdx.x];
    const int ca = ew & 0xF, cb = ew >> 4;
    static_assert((SY + 1) * RD + DW < (1 << 10), "corner offsets fit 10 bits");
    const uint32_t oa = (uint32_t)((mc_dz(ca) * SY + mc_dy(ca)) * RD + mc_dx(ca) * DW);
    const uint32_t ob = (uint32_t)((mc_dz(cb) * SY + mc_dy(cb)) * RD + mc_dx(cb) * DW);
    sm.le[threadIdx.x] = oa | (ob << 10) |
                         ((uint32_t)(mc_dx(ca) | (mc_dy(ca) << 1) | (mc_dz(ca) << 2)) << 20) |
                         ((uint32_t)(mc_dx(cb) | (mc_dy(cb) << 1) | (mc_dz(cb) << 2)) << 23);
  }

  auto tile = [&](const VolArgs& a, const int64_t t) {
    const T* __restrict__ f = (const T*)a.f;
    Out* __restrict__ verts = (Out*)a.verts;
    int64_t* __restrict__ faces = a.faces;
    const Geom g = a.g;
    // x fastest: neighbouring tiles along x (which share the 128-byte lines
    // at their ends) run together and meet in L2
    const int ib = (int)(t % a.nti) * TX;
    const int64_t rest = t / a.nti;
    const int jb = (int)(rest % a.ntj) * kMc5CY;
    const int s = (int)(rest / a.ntj);
    const int zs = s * g.zseg, ze = min(zs + g.zseg, g.nz - 1);
    const T thr = (T)a.thr;
    const M iso = (M)a.iso;
    const int cxm = CXM >= 0 ? CXM : a.cxmask;
    const bool cx32 = cxm & 1, cy32 = (cxm >> 1) & 1, cz32 = (cxm >> 2) & 1;
    const double* zp = a.coords + g.nx + g.ny;
    const int64_t vbase = a.sz[0], totv = a.sz[1], fbase = a.sz[2], totf = a.sz[3];
    const int64_t nelem = (int64_t)g.nx * g.ny * g.nz;
    const int64_t pstride = (int64_t)g.nx * g.ny;
    for (int c = threadIdx.x; c < SX; c += kBlock) sm.cx[c] = a.coords[min(ib + c, g.nx - 1)];
    if (threadIdx.x < SY) sm.cy[threadIdx.x] = a.coords[g.nx + min(jb + (int)threadIdx.x, g.ny - 1)];

    // this lane's row (ib + lane, jb + wv): output offsets, carried chunk to chunk
    const bool rowok = lane < TX && ib + lane < g.nx - 1 && jb + wv < g.ny - 1;
    int64_t rbv = 0, rbf = 0;
    uint64_t zm = 0;
    if (rowok) {
      const int64_t u = a.ubase + ((int64_t)(ib + lane) * (g.ny - 1) + jb + wv) * g.nzseg + s;
      const uint64_t zu = zmask_at<T>(a, ib + lane, jb + wv, s);  // beside the offsets
      rbv = a.voff[u] - vbase;
      rbf = a.foff[u] - fbase;
      zm = a.cnt[u] != 0 ? zu : 0ull;
    }
    // the z blocks where some row of the tile may hold an active cell: the
    // chunks of the other blocks write nothing and are never fetched, and a
    // tile with none is skipped whole (sparse fields have many of both)
    bool any;
    const uint64_t tmask = tile_zmask(zm, zm != 0, sm.tzm, &any);
    if (tmask == 0) return;
    const int zsh = zmask_shift(g.zseg);
    const int z0 = zs + (__builtin_ctzll(tmask) << zsh);

    // DMA layout (as in the emit kernel with lanes along z): per-lane byte
    // offsets relative to the chunk's first sample are fixed.
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
      const int64_t start = (int64_t)zc * pstride + (int64_t)jb * g.nx + ib;
      const int64_t left = (nelem - start) * (int64_t)sizeof(T);
      const __amdgpu_buffer_rsrc_t rs =
          buffer_rsrc(f + start, (uint32_t)(left < 0x7FFFFFFF ? left : 0x7FFFFFFF));
  #pragma unroll
      for (int m = 0; m < NIW; ++m) {
        const int ins = wv + m * kWaves;
        if (dok[m]) dma_to_lds<VEC ? 16 : 4>(rs, sm.smp[buf] + ins * (VEC ? 256 : kWave), doff[m]);
      }
      constexpr int NZ = (2 * NP + kWave - 1) / kWave;
      uint32_t* zd = (uint32_t*)sm.cz[buf];
      for (int ins = wv; ins < NZ; ins += kWaves) {
        const int p = ins * kWave + lane;
        if (p < 2 * NP) {
          const double* src = zp + min(zc + (p >> 1), g.nz - 1);
          __builtin_amdgcn_global_load_lds((const void*)((const char*)src + 4 * (p & 1)),
                                           (void*)(zd + ins * kWave), 4, 0, 0);
        }
      }
    };

    uint32_t* list = sm.list[wv];
    uint16_t* lpf = sm.lpf[wv];
    uint64_t* vbits = sm.vbits[wv];
    uint64_t* fbits = sm.fbits[wv];
    int64_t* dvw = sm.dv[wv];
    int64_t* dfw = sm.df[wv];
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

      // classify: the face nibble of plane p for cell (lane, wv) in natural
      // order (x,y) (x+1,y) (x,y+1) (x+1,y+1); the x+1 bits come from lane+1
      uint32_t nib[NP];
  #pragma unroll
      for (int p = 0; p < NP; ++p) {
        const T v0 = *(const T*)(smp + (p * SY + wv) * RD + lane * DW);
        const T v1 = *(const T*)(smp + (p * SY + wv + 1) * RD + lane * DW);
        const uint32_t w = (uint32_t)(v0 < thr) | ((uint32_t)(v1 < thr) << 2);
        nib[p] = w | (next_lane(w, 0u) << 1);
      }
      const int zn = min(ZC, ze - zc);  // cells of this chunk along z
      uint32_t cs[ZC], pk[ZC];
      uint32_t tot = 0;  // the row's vertices | faces << 16 in this chunk
  #pragma unroll
      for (int z = 0; z < ZC; ++z) {
        cs[z] = (rowok && z < zn) ? (nib[z] | (nib[z + 1] << 4)) : 0u;
        pk[z] = sm.cnt[cs[z]];
        tot += pk[z];
      }

      int nl = 0;          // entries of the wave's queue
      uint32_t V = 0, F = 0;  // vertex and face slots of the wave's queue
      auto flush = [&]() {
        // slot o belongs to the last entry starting at or before o
        int before = 0;
        for (uint32_t o0 = 0; o0 < V; o0 += kWave) {
          const uint32_t o = o0 + lane;
          const uint64_t word = vbits[o0 >> 6];
          const uint64_t upto = word & (~0ull >> (kWave - 1 - lane));
          const int k = before + __popcll(upto) - 1;
          before += __popcll(word);
          if (o >= V) break;
          const uint32_t e = list[k];
          const int ci = (int)(e & 63), c = (int)((e >> 6) & 15);
          const uint32_t ecs = (e >> 10) & 0xFF, sv = e >> 18;
          const uint32_t q = o - sv;
          const uint32_t ed = (uint32_t)(sm.lv[ecs] >> (4 + 4 * q)) & 0xF;
          const uint32_t ew = sm.le[ed];
          const uint32_t* cell = smp + (c * SY + wv) * RD + ci * DW;
          const T va = *(const T*)(cell + (ew & 0x3FF)), vb = *(const T*)(cell + ((ew >> 10) & 0x3FF));
          const M mu = (iso - (M)va) / (M)(vb - va);
          const double px = lerp_axis<M>(cx32, mu, sm.cx[ci + ((ew >> 20) & 1)], sm.cx[ci + ((ew >> 23) & 1)]);
          const double py = lerp_axis<M>(cy32, mu, sm.cy[wv + ((ew >> 21) & 1)], sm.cy[wv + ((ew >> 24) & 1)]);
          const double pz = lerp_axis<M>(cz32, mu, czs[c + ((ew >> 22) & 1)], czs[c + ((ew >> 25) & 1)]);
          const int64_t vi = dvw[ci] + o;
          if (vi < totv) {
            // plain stores: partial lines of a row's run meet the next chunk's
            // in L2 (streaming stores measured slower: they bypass that merge)
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
          const uint32_t e = list[k];
          const int ci = (int)(e & 63);
          const uint32_t ecs = (e >> 10) & 0xFF, sv = e >> 18;
          const uint32_t q = o - lpf[k];
          const uint64_t w = sm.lt[ecs] >> (3 + 12 * q);
          const int64_t vb = dvw[ci] + sv + 1;  // the cell's first vertex, 1-based
          const int64_t fi = dfw[ci] + o;
          if (fi < totf) {
            int64_t* out = faces + fi * 3;
            out[0] = vb + (int64_t)(w & 0xF);
            out[1] = vb + (int64_t)((w >> 4) & 0xF);
            out[2] = vb + (int64_t)((w >> 8) & 0xF);
          }
        }
        wave_sync();
        for (uint32_t q = lane; q < (V + 63) / 64; q += kWave) vbits[q] = 0;
        for (uint32_t q = lane; q < (F + 63) / 64; q += kWave) fbits[q] = 0;
      };

      // queue the cells z in [za, zb) of every row: rows one after another,
      // each row's cells in z order, which is the order of their outputs
      auto append = [&](int za, int zb, uint32_t rtot, uint32_t run0) {
        uint32_t ne = 0;
  #pragma unroll
        for (int z = 0; z < ZC; ++z) ne += (z >= za && z < zb && pk[z] != 0) ? 1u : 0u;
        const uint32_t incv = wave_scan_dpp(rtot), ince = wave_scan_dpp(ne);
        const uint32_t sumv = (uint32_t)__builtin_amdgcn_readlane((int)incv, kWave - 1);
        const uint32_t sume = (uint32_t)__builtin_amdgcn_readlane((int)ince, kWave - 1);
        const uint32_t bv = V + ((incv - rtot) & 0xFFFF), bf = F + ((incv - rtot) >> 16);
        uint32_t k = (uint32_t)nl + ince - ne;
        // output index of queue slot o of this row: o + d
        if (ne) {
          dvw[lane] = rbv + (int64_t)(run0 & 0xFFFF) - (int64_t)bv;
          dfw[lane] = rbf + (int64_t)(run0 >> 16) - (int64_t)bf;
        }
        uint32_t run = 0;
  #pragma unroll
        for (int z = 0; z < ZC; ++z) {
          const bool in = z >= za && z < zb;
          if (in && pk[z] != 0) {
            const uint32_t sv = bv + (run & 0xFFFF), sf = bf + (run >> 16);
            list[k] = (uint32_t)lane | ((uint32_t)z << 6) | (cs[z] << 10) | (sv << 18);
            lpf[k] = (uint16_t)sf;
            atomicOr((unsigned long long*)&vbits[sv >> 6], 1ull << (sv & 63));
            atomicOr((unsigned long long*)&fbits[sf >> 6], 1ull << (sf & 63));
            ++k;
          }
          if (in) run += pk[z];
        }
        nl += (int)sume;
        V += sumv & 0xFFFF;
        F += sumv >> 16;
      };

      // this wave's share of the next brick must land before the barrier that
      // ends the chunk; wait now, so the wait does not also cover the stores
      // the queue is about to issue
      __builtin_amdgcn_s_waitcnt(0);
      const uint32_t ttot = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_dpp(tot), kWave - 1);
      uint32_t nact = 0;
  #pragma unroll
      for (int z = 0; z < ZC; ++z) nact += pk[z] != 0 ? 1u : 0u;
      const uint32_t tact = (uint32_t)__builtin_amdgcn_readlane((int)wave_scan_dpp(nact), kWave - 1);
      // the whole chunk in one go when it fits the queue, else (a dense
      // chunk) one z step at a time
      const bool whole = (ttot & 0xFFFF) <= (uint32_t)CAPV && (ttot >> 16) <= (uint32_t)CAPF &&
                         tact <= (uint32_t)CAPE;
      uint32_t run0 = 0;
      for (int za = 0; ttot != 0 && za < zn;) {
        const int zb = whole ? ZC : za + 1;
        uint32_t part = 0;
  #pragma unroll
        for (int q = 0; q < ZC; ++q) part += (q >= za && q < zb) ? pk[q] : 0u;
        append(za, zb, part, run0);
        run0 += part;
        wave_sync();
        flush();
        wave_sync();
        nl = 0;
        V = 0;
        F = 0;
        za = zb;
      }
      rbv += tot & 0xFFFF;
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

template <typename T, typename M, typename Out, int ZC, bool VEC, int CXM, int NBUF>
__global__ __launch_bounds__(kBlock) void k_emit_mc5(VolArgs a) {
  emit_mc5_run<T, M, Out, ZC, VEC, CXM, NBUF>(
      OneTile{&a, xcd_tile(blockIdx.x, a.ntiles), false});
}

// Batched, persistent: the launch's blocks take the tiles of its volumes
// from a work queue (TileQueue).
template <typename T, typename M, typename Out, int ZC, bool VEC, int CXM, int NBUF>
__global__ __launch_bounds__(kBlock) void k_emit_mc5_b(const VolArgs* __restrict__ vols,
                                                       const int32_t* __restrict__ emap,
                                                       int64_t gbase, int64_t ntiles,
                                                       uint32_t* __restrict__ ctr) {
  emit_mc5_run<T, M, Out, ZC, VEC, CXM, NBUF>(
      TileQueue<true>{vols, emap, gbase, ntiles, ctr, 0});
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

template <typename T>
static constexpr int mc5_zc() { return sizeof(T) == 4 ? 16 : 8; }
// bricks per block: 2 for the full-depth chunks (the next chunk's brick
// streams in during this one), 1 for the half-depth ones of family 3, which
// then fit five blocks per CU: on the sparse 256^3 gyroid 0.056 ms against
// 0.062 double-buffered (and 0.073 full-depth), while the dense 1024^3 field
// keeps the double-buffered full depth (4.67 ms; single-buffered 4.96)
template <typename T, int ZC>
static constexpr int mc5_nbuf() { return ZC == mc5_zc<T>() / 2 ? 1 : 2; }

// Half-depth bricks (family 3): the z chunk of the all-Float32 and
// all-Float64 instantiations with Float64 axes is halved (a smaller LDS
// brick, more blocks per CU).
int mc5_variant(const Plan& p, VolArgs* a) {
  const bool half = p.family == 3 && p.T == p.M && p.M == p.Out && a->cxmask == 0;
  const int zc = (p.T == K_F32 ? mc5_zc<float>() : mc5_zc<double>()) / (half ? 2 : 1);
  const int64_t esz = p.T == K_F32 ? 4 : 8;
  // the brick's 32-bit DMA offsets reach plane zc, row 4, sample 63
  if (((int64_t)zc * a->g.nx * a->g.ny + 4ll * a->g.nx + 64) * esz >= (1ll << 31)) return -1;
  a->nti = (a->g.nx - 1 + kMc5TX - 1) / kMc5TX;
  a->ntj = (a->g.ny - 1 + kMc5CY - 1) / kMc5CY;
  a->ntiles = (a->g.nx > 1 && a->g.ny > 1) ? (int64_t)a->nti * a->ntj * a->g.nzseg : 0;
  const bool vec = ((uintptr_t)a->f % 16 == 0) && (((int64_t)a->g.nx * esz) % 16 == 0);
  const int cxclass = a->cxmask == 0 ? 1 : a->cxmask == 7 ? 2 : 0;
  return (vec ? 0 : 1) | (cxclass << 1) | (half ? 8 : 0);
}

template <typename T, typename M, typename Out, int CXM, int ZC = mc5_zc<T>()>
static hipError_t mc5_z(bool vec, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                        int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  constexpr int NB = mc5_nbuf<T, ZC>();
  if (vols) {
    const auto k = vec ? k_emit_mc5_b<T, M, Out, ZC, true, CXM, NB>
                       : k_emit_mc5_b<T, M, Out, ZC, false, CXM, NB>;
    hipLaunchKernelGGL(k, dim3(persistent_grid((const void*)k, nblocks)), dim3(kBlock), 0, st,
                       vols, emap, gbase, nblocks, ctr);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_emit_mc5<T, M, Out, ZC, true, CXM, NB>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_emit_mc5<T, M, Out, ZC, false, CXM, NB>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

template <typename T, typename M, typename Out>
static hipError_t mc5_t(int variant, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                        int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  const bool vec = (variant & 1) == 0;
  if constexpr (std::is_same<T, M>::value && std::is_same<M, Out>::value) {
    if (variant & 8)
      return mc5_z<T, M, Out, 0, mc5_zc<T>() / 2>(vec, a, vols, emap, gbase, nblocks, ctr, st);
  }
  switch ((variant >> 1) & 3) {
    case 1: return mc5_z<T, M, Out, 0>(vec, a, vols, emap, gbase, nblocks, ctr, st);
    case 2: return mc5_z<T, M, Out, 7>(vec, a, vols, emap, gbase, nblocks, ctr, st);
    default: return mc5_z<T, M, Out, -1>(vec, a, vols, emap, gbase, nblocks, ctr, st);
  }
}

static hipError_t mc5_any(const Plan& p, int variant, const VolArgs& a, const VolArgs* vols,
                          const int32_t* emap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                          hipStream_t st) {
#define MSH_ARGS variant, a, vols, emap, gbase, nblocks, ctr, st
  if (p.T == K_F32 && p.M == K_F32 && p.Out == K_F32) return mc5_t<float, float, float>(MSH_ARGS);
  if (p.T == K_F32 && p.M == K_F32) return mc5_t<float, float, double>(MSH_ARGS);
  if (p.T == K_F32) return mc5_t<float, double, double>(MSH_ARGS);
  return mc5_t<double, double, double>(MSH_ARGS);
#undef MSH_ARGS
}

hipError_t launch_emit_mc5(const Plan& p, int variant, const VolArgs& a, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
  return mc5_any(p, variant, a, nullptr, nullptr, 0, 0, nullptr, st);
}

hipError_t launch_emit_mc5_batch(const Plan& p, int variant, const VolArgs* vols,
                                 const int32_t* emap, int64_t gbase, int64_t nblocks,
                                 uint32_t* ctr, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  return mc5_any(p, variant, VolArgs{}, vols, emap, gbase, nblocks, ctr, st);
}

}  // namespace msh
