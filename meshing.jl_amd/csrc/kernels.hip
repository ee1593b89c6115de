// kernels.hip -- gfx950 kernels of the isosurface path.
//
// Pipeline per extraction (all on one stream):
//   k_coords   LinRange coordinates of the three axes (Julia lerpi).
//   k_count    classify every cell and count, per z-row segment ("unit"),
//              the vertices and faces it will emit.  Lanes run along x
//              (coalesced HBM reads); each lane marches z, keeping the
//              previous sample plane in registers and taking the x+1
//              neighbour from the next lane.
//   k_scan_*   exclusive prefix of the unit counts in the reference's
//              emission order (x outermost, z innermost), reduce-then-scan.
//   k_emit_*   re-read the field brick by brick into LDS, run each z-row
//              chunk of 64 cells as one wavefront (lanes along z, the
//              emission order), wave-scan the per-cell counts and write the
//              chunk's vertices and faces slot-parallel: lane s writes
//              output s, so stores are contiguous.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <map>
#include <mutex>
#include <utility>

#include "internal.h"
#include "device.h"

#define MSH_LUT static __constant__
#include "luts.h"

namespace msh {
namespace {

// ---------------------------------------------------------------------------
// LinRange coordinates: xp[i] = CX((1-t)*a + t*b), t = (i-1)/max(1, n-1)
// (Julia Base lerpi; used at src/marching_cubes.jl:36-38).
__device__ __forceinline__ double lerpi(int j, int n, double a, double b, bool f32) {
  const int d = n - 1 > 1 ? n - 1 : 1;
  const double t = (double)j / (double)d;
  const double v = (1.0 - t) * a + t * b;
  return f32 ? (double)(float)v : v;
}

// coordinate i of the three axes laid end to end (x, then y, then z)
__device__ __forceinline__ double coord_at(const CoordArgs& c, int i) {
  int axis = 0, n = c.nx, j = i;
  if (i >= c.nx + c.ny) {
    axis = 2; n = c.nz; j = i - c.nx - c.ny;
  } else if (i >= c.nx) {
    axis = 1; n = c.ny; j = i - c.nx;
  }
  return lerpi(j, n, c.a[axis], c.b[axis], (c.mask >> axis) & 1);
}

__global__ void k_coords(CoordArgs c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c.nx + c.ny + c.nz) c.out[i] = coord_at(c, i);
}

// Batched: one block per volume.
__global__ void k_coords_b(const VolArgs* __restrict__ vols) {
  const VolArgs& a = vols[blockIdx.x];
  const int nx = a.g.nx, ny = a.g.ny, nz = a.g.nz;
  for (int i = threadIdx.x; i < nx + ny + nz; i += blockDim.x) {
    int axis = 0, n = nx, j = i;
    if (i >= nx + ny) {
      axis = 2; n = nz; j = i - nx - ny;
    } else if (i >= nx) {
      axis = 1; n = ny; j = i - nx;
    }
    a.coords[i] = lerpi(j, n, a.ca[axis], a.cb[axis], (a.cxmask >> axis) & 1);
  }
}

// Block -> volume maps of a batch: one block per volume fills its ranges.
__global__ void k_fill_maps(const VolArgs* __restrict__ vols, int32_t* __restrict__ cmap,
                            int32_t* __restrict__ emap) {
  const int v = blockIdx.x;
  const VolArgs& a = vols[v];
  for (int64_t i = threadIdx.x; i < a.cntiles; i += blockDim.x) cmap[a.cblk0 + i] = v;
  for (int64_t i = threadIdx.x; i < a.ntiles; i += blockDim.x) emap[a.eblk0 + i] = v;
}

// Per-volume sizes of a batch from the scan over all volumes' units (volume
// v's units start at ubase and volumes are stored in unit order):
// (first vertex, nverts, first face, nfaces).
__global__ void k_sizes(const VolArgs* __restrict__ vols, int n, const int64_t* __restrict__ voff,
                        const int64_t* __restrict__ foff, const int64_t* __restrict__ totals,
                        int64_t units, int64_t* __restrict__ sizes) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const int64_t ub = vols[v].ubase;
  const int64_t ue = v + 1 < n ? vols[v + 1].ubase : units;
  const int64_t vb = ub < units ? voff[ub] : totals[1], fb = ub < units ? foff[ub] : totals[3];
  const int64_t ve = ue < units ? voff[ue] : totals[1], fe = ue < units ? foff[ue] : totals[3];
  sizes[4 * v + 0] = vb;
  sizes[4 * v + 1] = ve - vb;
  sizes[4 * v + 2] = fb;
  sizes[4 * v + 3] = fe - fb;
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Scan: exclusive prefix of the unit counts (vertices in bits 0-15 of a
// count, faces in bits 16-31) over tiles of kScanTile units, in three
// launches that never wait on one another's blocks: k_scan_reduce sums each
// tile, k_scan_tiles (one block) turns the tile sums into tile offsets and
// the totals, k_scan_apply scans each tile from its offset (up to
// kScanFuseTiles tiles the last two are one launch, k_scan_apply_sums).  A single-pass
// scan with decoupled look-back took 1.1 ms over the 10.9 M units of a
// 512-volume batch: every tile's inclusive prefix is a round trip through
// memory that the next tiles wait on.

// This thread's kScanItems counts and their packed sum (vertices in the low
// 32 bits, faces in the high 32 bits; a tile's sums fit both halves).
__device__ __forceinline__ uint64_t scan_load(const uint32_t* __restrict__ cnt, int64_t units,
                                              int64_t base, uint32_t (&c)[kScanItems]) {
  uint64_t sum = 0;
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    c[q] = base + q < units ? cnt[base + q] : 0u;
    sum += (uint64_t)(c[q] & 0xFFFF) | ((uint64_t)(c[q] >> 16) << 32);
  }
  return sum;
}

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* __restrict__ cnt,
                                                        int64_t units,
                                                        uint64_t* __restrict__ tilesum) {
  __shared__ uint64_t s_w[kWaves];
  uint32_t c[kScanItems];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const uint64_t incl = wave_incl_scan64(scan_load(cnt, units, base, c));
  if (lane_id() == kWave - 1) s_w[wave_id()] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t agg = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) agg += s_w[q];
    tilesum[blockIdx.x] = agg;
  }
}

constexpr int kScanTilesBlock = 1024;

__global__ __launch_bounds__(kScanTilesBlock) void k_scan_tiles(
    const uint64_t* __restrict__ tilesum, int64_t ntiles, int64_t* __restrict__ tileoff,
    int64_t* __restrict__ totals, int64_t* __restrict__ host, CoordArgs coords) {
  // the volume's coordinate tables, when asked for: one launch fewer
  if (coords.out)
    for (int i = threadIdx.x; i < coords.nx + coords.ny + coords.nz; i += kScanTilesBlock)
      coords.out[i] = coord_at(coords, i);
  __shared__ int64_t s_v[kScanTilesBlock / kWave], s_f[kScanTilesBlock / kWave];
  const int t = threadIdx.x, lane = lane_id(), wv = t / kWave;
  const int64_t per = (ntiles + kScanTilesBlock - 1) / kScanTilesBlock;
  const int64_t lo = t * per, hi = lo + per < ntiles ? lo + per : ntiles;
  int64_t sv = 0, sf = 0;  // int64: the sums over many tiles outgrow 32 bits
  for (int64_t k = lo; k < hi; ++k) {
    sv += (int64_t)(tilesum[k] & 0xFFFFFFFFu);
    sf += (int64_t)(tilesum[k] >> 32);
  }
  const int64_t iv = (int64_t)wave_incl_scan64((uint64_t)sv);
  const int64_t jf = (int64_t)wave_incl_scan64((uint64_t)sf);
  if (lane == kWave - 1) {
    s_v[wv] = iv;
    s_f[wv] = jf;
  }
  __syncthreads();
  int64_t bv = 0, bf = 0;
  for (int q = 0; q < wv; ++q) {
    bv += s_v[q];
    bf += s_f[q];
  }
  int64_t rv = bv + iv - sv, rf = bf + jf - sf;
  for (int64_t k = lo; k < hi; ++k) {
    tileoff[2 * k] = rv;
    tileoff[2 * k + 1] = rf;
    rv += (int64_t)(tilesum[k] & 0xFFFFFFFFu);
    rf += (int64_t)(tilesum[k] >> 32);
  }
  if (t == kScanTilesBlock - 1) {
    totals[0] = 0;
    totals[1] = bv + iv;
    totals[2] = 0;
    totals[3] = bf + jf;
    if (host) {
      // the sizes straight into the caller's pinned host words (no copy
      // after the scan), made visible to the host before the kernel ends
      host[1] = bv + iv;
      host[3] = bf + jf;
      __threadfence_system();
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* __restrict__ cnt,
                                                       int64_t units,
                                                       const int64_t* __restrict__ tileoff,
                                                       int64_t* __restrict__ voff,
                                                       int64_t* __restrict__ foff) {
  __shared__ uint64_t s_w[kWaves];
  uint32_t c[kScanItems];
  const int wv = wave_id();
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const uint64_t sum = scan_load(cnt, units, base, c);
  const uint64_t wincl = wave_incl_scan64(sum);
  if (lane_id() == kWave - 1) s_w[wv] = wincl;
  __syncthreads();
  uint64_t before = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) before += q < wv ? s_w[q] : 0ull;
  const uint64_t texcl = before + wincl - sum;  // within the tile: both halves fit
  int64_t rv = tileoff[2 * blockIdx.x] + (int64_t)(texcl & 0xFFFFFFFFu);
  int64_t rf = tileoff[2 * blockIdx.x + 1] + (int64_t)(texcl >> 32);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < units) {
      voff[base + q] = rv;
      foff[base + q] = rf;
    }
    rv += c[q] & 0xFFFF;
    rf += c[q] >> 16;
  }
}

// Up to kScanFuseTiles tiles, k_scan_tiles and k_scan_apply are one launch:
// each block adds up the sums of the tiles before it itself (at most a few
// KB of reads), and the last block also writes the totals.  One launch
// fewer per count phase; past that size the one-block tile scan is
// cheaper than every block reading the earlier sums.
constexpr int64_t kScanFuseTiles = 1024;

__global__ __launch_bounds__(kBlock) void k_scan_apply_sums(
    const uint32_t* __restrict__ cnt, int64_t units, const uint64_t* __restrict__ tilesum,
    int64_t* __restrict__ voff, int64_t* __restrict__ foff, int64_t* __restrict__ totals,
    int64_t* __restrict__ host, CoordArgs coords) {
  __shared__ uint64_t s_w[kWaves];
  __shared__ int64_t s_v[kWaves], s_f[kWaves];
  uint32_t c[kScanItems];
  const int wv = wave_id(), lane = lane_id();
  // this tile's offset: the sums of the tiles before it (int64: they
  // outgrow 32 bits over many tiles)
  int64_t pv = 0, pf = 0;
  for (int64_t k = threadIdx.x; k < (int64_t)blockIdx.x; k += kBlock) {
    pv += (int64_t)(tilesum[k] & 0xFFFFFFFFu);
    pf += (int64_t)(tilesum[k] >> 32);
  }
  const int64_t iv = (int64_t)wave_incl_scan64((uint64_t)pv);
  const int64_t jf = (int64_t)wave_incl_scan64((uint64_t)pf);
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  const uint64_t sum = scan_load(cnt, units, base, c);
  const uint64_t wincl = wave_incl_scan64(sum);
  if (lane == kWave - 1) {
    s_w[wv] = wincl;
    s_v[wv] = iv;
    s_f[wv] = jf;
  }
  __syncthreads();
  uint64_t before = 0, tile = 0;
  int64_t tv = 0, tf = 0;
#pragma unroll
  for (int q = 0; q < kWaves; ++q) {
    before += q < wv ? s_w[q] : 0ull;
    tile += s_w[q];
    tv += s_v[q];
    tf += s_f[q];
  }
  const uint64_t texcl = before + wincl - sum;  // within the tile: both halves fit
  int64_t rv = tv + (int64_t)(texcl & 0xFFFFFFFFu);
  int64_t rf = tf + (int64_t)(texcl >> 32);
#pragma unroll
  for (int q = 0; q < kScanItems; ++q) {
    if (base + q < units) {
      voff[base + q] = rv;
      foff[base + q] = rf;
    }
    rv += c[q] & 0xFFFF;
    rf += c[q] >> 16;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    const int64_t nv = tv + (int64_t)(tile & 0xFFFFFFFFu), nf = tf + (int64_t)(tile >> 32);
    totals[0] = 0;
    totals[1] = nv;
    totals[2] = 0;
    totals[3] = nf;
    if (host) {
      host[1] = nv;
      host[3] = nf;
      __threadfence_system();
    }
  }
  if (blockIdx.x == 0 && coords.out)
    for (int i = threadIdx.x; i < coords.nx + coords.ny + coords.nz; i += kBlock)
      coords.out[i] = coord_at(coords, i);
}

// ---------------------------------------------------------------------------
// Emit.
//
// One block = a brick of TX x TY cell rows (plus, for MT and SN, a one-cell
// halo on the low x and y sides); it walks the z segment of its tile in
// chunks of 64 cells.  Per chunk the samples of the brick are staged in LDS
// (plane stride padded to an odd number of words so lanes reading along z
// hit distinct banks); the next chunk's samples are already being fetched
// into registers while this one is processed.  Each wave takes cell rows
// (i, j): lane l is cell (i, j, zc+l), so the lanes follow the reference's
// emission order, one wave scan gives every cell's output slot, and the
// chunk's outputs of a row form one contiguous range per output array.
// Outputs are written slot-parallel (lane s writes output s of the row).
//
// MT and SN faces refer to vertices created by the neighbouring cells at
// -x, -y, -z.  Those cells' cases and in-row vertex prefixes are in LDS
// (halo rows, and the plane below the chunk), so a face's vertex id is
// (row base + prefix of the owning cell + rank of the edge in that cell).
template <int METHOD, typename T, int TX, int TY>
struct EmitCfg {
  static constexpr int H = METHOD == MC ? 0 : 1;       // low-side halo
  static constexpr int RX = TX + H, RY = TY + H;       // cell rows incl. halo
  static constexpr int NR = RX * RY;
  static constexpr int SX = RX + 1, SY = RY + 1;       // samples per plane
  static constexpr int NP = kWave + 1 + H;             // sample planes per chunk
  static constexpr int PS = (SX * SY) | 1;             // padded plane stride
  static constexpr int CAP = 256;                      // active-cell list per wave
};

template <int METHOD, typename T, int TX, int TY>
struct EmitSmem {
  using C = EmitCfg<METHOD, T, TX, TY>;
  T smp[2][C::PS * C::NP];          // double-buffered brick, filled by LDS DMA
  double cz[2][C::NP];
  double cx[C::SX], cy[C::SY];
  int64_t rbv[2][C::NR], rbf[2][C::NR];  // offsets of each row's first cell: this/next chunk
  uint8_t ccase[C::H ? C::NR * kWave : 1];
  uint16_t cpv[C::H ? C::NR * kWave : 1];
  uint8_t pcase[C::H ? C::NR : 1];  // case of the cell just below the chunk
  uint64_t lut_v[METHOD == MC ? 256 : 1];  // MC_CASE_VERTS
  uint64_t lut_t[METHOD == MC ? 256 : 1];  // MC_CASE_TRIS
  uint8_t lut_e[16];                       // MC_EDGE
  uint64_t list[kWaves][C::CAP];    // active cells: row | z<<8 | case<<14 | pv<<22 | pf<<32
};

template <int METHOD, typename T, typename M, typename A, typename Out, int TX, int TY>
__device__ __forceinline__ void emit_tile(const VolArgs& a, int64_t t) {
  using C = EmitCfg<METHOD, T, TX, TY>;
  constexpr int H = C::H, RX = C::RX, NR = C::NR, SX = C::SX, SY = C::SY, NP = C::NP,
                PS = C::PS, CAP = C::CAP;
  __shared__ EmitSmem<METHOD, T, TX, TY> sm;
  const T* __restrict__ f = (const T*)a.f;
  Out* __restrict__ verts = (Out*)a.verts;
  int64_t* __restrict__ faces = a.faces;
  const Geom g = a.g;
  const int lane = lane_id(), wv = wave_id();

  const int tj = (int)(t % a.ntj);
  const int64_t rest = t / a.ntj;
  const int ti = (int)(rest % a.nti);
  const int s = (int)(rest / a.nti);
  const int x0 = ti * TX - H, y0 = tj * TY - H;  // global cell index of brick row 0
  const int zs = s * g.zseg, ze = min(zs + g.zseg, g.nz - 1);
  const T thr = (T)a.thr;
  const M iso = (M)a.iso;
  const bool cx32 = a.cxmask & 1, cy32 = (a.cxmask >> 1) & 1, cz32 = (a.cxmask >> 2) & 1;
  const double* xp = a.coords;
  const double* yp = a.coords + g.nx;
  const double* zp = a.coords + g.nx + g.ny;
  const int64_t vbase = a.sz[0], totv = a.sz[1], fbase = a.sz[2], totf = a.sz[3];
  const int64_t pstride = (int64_t)g.nx * g.ny;

  if constexpr (METHOD == MC) {
    for (int q = threadIdx.x; q < 256; q += kBlock) {
      sm.lut_v[q] = MC_CASE_VERTS[q];
      sm.lut_t[q] = MC_CASE_TRIS[q];
    }
  }
  if (METHOD != MT && threadIdx.x < 12) sm.lut_e[threadIdx.x] = MC_EDGE[threadIdx.x];
  bool any = false;  // some row of the brick, halo rows included, has outputs
  for (int r = threadIdx.x; r < NR; r += kBlock) {
    const int i = x0 + r % RX, j = y0 + r / RX;
    int64_t bv = 0, bf = 0;
    if (i >= 0 && j >= 0 && i < g.nx - 1 && j < g.ny - 1) {
      const int64_t u = a.ubase + ((int64_t)i * (g.ny - 1) + j) * g.nzseg + s;
      bv = a.voff[u] - vbase;
      bf = a.foff[u] - fbase;
      any = any || a.cnt[u] != 0;
    }
    sm.rbv[0][r] = bv;
    sm.rbf[0][r] = bf;
  }
  // a tile whose rows the count found empty writes nothing: skip it
  // before fetching any of its bricks (sparse fields have many)
  if (!__syncthreads_or(any)) return;
  for (int q = threadIdx.x; q < SX; q += kBlock) sm.cx[q] = xp[min(max(x0 + q, 0), g.nx - 1)];
  for (int q = threadIdx.x; q < SY; q += kBlock) sm.cy[q] = yp[min(max(y0 + q, 0), g.ny - 1)];

  // DMA of the brick for the chunk starting at cell zc (planes zc-H ..
  // zc+64) straight into LDS buffer `buf`: each wave instruction fills 64
  // consecutive dwords; lanes landing on plane padding load a harmless
  // in-bounds dword.  No registers hold the data, so the next chunk's
  // loads stay in flight while this chunk is processed.
  auto fetch = [&](int zc, int buf) {
    constexpr int DW = (int)(sizeof(T) / 4);
    constexpr int NDW = PS * NP * DW;
    constexpr int NI = (NDW + kWave - 1) / kWave;
    uint32_t* dst = (uint32_t*)sm.smp[buf];
#pragma unroll 4
    for (int ins = wv; ins < NI; ins += kWaves) {
      const int p = ins * kWave + lane;
      if (p < NDW) {
        const int e = p / DW, half = p - e * DW;
        const int z = e / PS, rem = e - z * PS;
        const char* src = (const char*)f;
        if (rem < SX * SY) {
          const int y = rem / SX, x = rem - y * SX;
          const int gx = min(max(x0 + x, 0), g.nx - 1), gy = min(max(y0 + y, 0), g.ny - 1),
                    gz = min(max(zc - H + z, 0), g.nz - 1);
          src = (const char*)(f + ((int64_t)gz * pstride + (int64_t)gy * g.nx + gx)) + 4 * half;
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (void*)(dst + ins * kWave), 4, 0, 0);
      }
    }
    constexpr int NZ = (NP * 2 + kWave - 1) / kWave;  // cz: NP doubles
    uint32_t* zd = (uint32_t*)sm.cz[buf];
    for (int ins = wv; ins < NZ; ins += kWaves) {
      const int p = ins * kWave + lane;
      if (p < NP * 2) {
        const int e = p >> 1, half = p & 1;
        const double* src = zp + min(max(zc - H + e, 0), g.nz - 1);
        __builtin_amdgcn_global_load_lds((const void*)((const char*)src + 4 * half),
                                         (void*)(zd + ins * kWave), 4, 0, 0);
      }
    }
  };

  uint64_t* list = sm.list[wv];

  if (zs < ze) fetch(zs, 0);
  __builtin_amdgcn_s_waitcnt(0);  // this wave's share of the first brick landed
  lds_barrier();
  int buf = 0;
  for (int zc = zs; zc < ze; zc += kWave, buf ^= 1) {
    // Buffer buf^1 was last read in the previous chunk, which every wave
    // has left (barrier at its end): refill it for the next chunk now.
    if (zc + kWave < ze) fetch(zc + kWave, buf ^ 1);
    const T* smp = sm.smp[buf];
    const double* czs = sm.cz[buf];
    const int64_t* rbv = sm.rbv[buf];
    const int64_t* rbf = sm.rbf[buf];
    int64_t* rbv_next = sm.rbv[buf ^ 1];
    int64_t* rbf_next = sm.rbf[buf ^ 1];
    const int kc = zc + lane;  // this lane's cell z
    const bool zvalid = kc < ze;

    if constexpr (H) {
      // Phase 1: classify every row of the brick, halo rows included, and
      // keep case + in-row vertex prefix for the neighbour lookups.
      for (int r = wv; r < NR; r += kWaves) {
        const int ci = r % RX, cj = r / RX;
        const int gi = x0 + ci, gj = y0 + cj;
        uint32_t cs = 0, pc = 0, nv = 0;
        if (gi >= 0 && gj >= 0 && gi < g.nx - 1 && gj < g.ny - 1) {
          const T* p = smp + cj * SX + ci;
          const uint32_t lo = face_code<METHOD, T, SX>(p + (lane + 1) * PS, thr);
          const uint32_t top = face_code<METHOD, T, SX>(p + (kWave + 1) * PS, thr);
          const uint32_t below = face_code<METHOD, T, SX>(p, thr);
          const uint32_t hi = next_lane(lo, top);
          cs = zvalid ? (lo | (hi << 4)) : 0u;
          const uint32_t lo0 = (uint32_t)__builtin_amdgcn_readlane((int)lo, 0);
          pc = zc > 0 ? (below | (lo0 << 4)) : 0u;
          if (METHOD == MT) {
            const int cls = (gi == 0) | ((gj == 0) << 1) | ((kc == 0) << 2);
            nv = MT_NOWN[cls * 256 + cs];
          } else {
            nv = (cs != 0 && cs != 255) ? 1u : 0u;
          }
        }
        const uint32_t inc = wave_scan_dpp(nv);
        sm.ccase[r * kWave + lane] = (uint8_t)cs;
        sm.cpv[r * kWave + lane] = (uint16_t)(inc - nv);
        if (lane == 0) {
          sm.pcase[r] = (uint8_t)pc;
          rbv_next[r] = rbv[r] + (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
        }
      }
      lds_barrier();
    }

    // Emission of one active cell by one lane.  Entry: row | z<<8 |
    // case<<14 | in-row vertex prefix<<22 | in-row face prefix<<32.
    auto emit_cell = [&](uint64_t ent) {
      const int r = (int)(ent & 0xFF), c = (int)((ent >> 8) & 63);
      const uint32_t cs = (uint32_t)((ent >> 14) & 0xFF);
      const uint32_t pv = (uint32_t)((ent >> 22) & 0x3FF), pf = (uint32_t)((ent >> 32) & 0x3FF);
      const int ci = r % RX, cj = r / RX;
      const int gi = x0 + ci, gj = y0 + cj, gk = zc + c;
      const int64_t vb = rbv[r] + pv;  // this cell's first vertex (0-based)
      const int64_t fb = rbf[r] + pf;  // this cell's first face
      const T* pc = smp + (c + H) * PS + cj * SX + ci;  // corner (0,0,0) of the cell
      if constexpr (METHOD == MC) {
        const uint64_t vw = sm.lut_v[cs];
        const int nv = (int)(vw & 0xF);
        for (int q = 0; q < nv; ++q) {
          const uint32_t e = (uint32_t)(vw >> (4 + 4 * q)) & 0xF;
          const uint32_t ew = sm.lut_e[e];
          const int ca = ew & 0xF, cb = ew >> 4;
          const T va = pc[mc_dz(ca) * PS + mc_dy(ca) * SX + mc_dx(ca)];
          const T vbv = pc[mc_dz(cb) * PS + mc_dy(cb) * SX + mc_dx(cb)];
          const M mu = (iso - (M)va) / (M)(vbv - va);
          const double px = lerp_axis<M>(cx32, mu, sm.cx[ci + mc_dx(ca)], sm.cx[ci + mc_dx(cb)]);
          const double py = lerp_axis<M>(cy32, mu, sm.cy[cj + mc_dy(ca)], sm.cy[cj + mc_dy(cb)]);
          const double pz = lerp_axis<M>(cz32, mu, czs[c + mc_dz(ca)], czs[c + mc_dz(cb)]);
          if (vb + q < totv) {
            Out* out = verts + (vb + q) * 3;
            out[0] = (Out)px;
            out[1] = (Out)py;
            out[2] = (Out)pz;
          }
        }
        const uint64_t tw = sm.lut_t[cs];
        const int nf = (int)(tw & 0x7);
        for (int q = 0; q < nf; ++q) {
          const uint64_t w = tw >> (3 + 12 * q);
          if (fb + q < totf) {
            int64_t* out = faces + (fb + q) * 3;
            out[0] = vb + 1 + (int64_t)(w & 0xF);
            out[1] = vb + 1 + (int64_t)((w >> 4) & 0xF);
            out[2] = vb + 1 + (int64_t)((w >> 8) & 0xF);
          }
        }
      } else if constexpr (METHOD == MT) {
        const int cls = (gi == 0) | ((gj == 0) << 1) | ((gk == 0) << 2);
        const int nv = MT_NOWN[cls * 256 + cs];
        for (int q = 0; q < nv; ++q) {
          const uint32_t e = MT_OWN_EDGES[(cls * 256 + cs) * 20 + q];
          const uint32_t ew = MT_EDGE[e];
          const int ca = ew & 0xF, cb = ew >> 4;
          const T sv = pc[mt_dz(ca) * PS + mt_dy(ca) * SX + mt_dx(ca)];
          const T tgv = pc[mt_dz(cb) * PS + mt_dy(cb) * SX + mt_dx(cb)];
          const M rr = (iso - (M)sv) / (M)(tgv - sv);
          const A aw = jl_min<A>(jl_max<A>((A)rr, (A)a.eps), (A)a.ome);
          const A bw = (A)1 - aw;
          const A wx = (A)mt_dx(ca) * bw + (A)mt_dx(cb) * aw;
          const A wy = (A)mt_dy(ca) * bw + (A)mt_dy(cb) * aw;
          const A wz = (A)mt_dz(ca) * bw + (A)mt_dz(cb) * aw;
          const double px = mt_axis<A>(cx32, wx, sm.cx[ci], sm.cx[ci + 1]);
          const double py = mt_axis<A>(cy32, wy, sm.cy[cj], sm.cy[cj + 1]);
          const double pz = mt_axis<A>(cz32, wz, czs[c + H], czs[c + H + 1]);
          if (vb + q < totv) {
            Out* out = verts + (vb + q) * 3;
            out[0] = (Out)px;
            out[1] = (Out)py;
            out[2] = (Out)pz;
          }
        }
        const int nf = MT_NF[cs];
        for (int q = 0; q < nf; ++q) {
          int64_t id[3];
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            const uint32_t e = MT_FACE[cs * 36 + 3 * q + m];
            const int qc = MT_EDGE[e] >> 4;  // end corner
            const int dx = (mt_dx(qc) == 0 && gi > 0), dy = (mt_dy(qc) == 0 && gj > 0),
                      dz = (mt_dz(qc) == 0 && gk > 0);
            const uint32_t e2 = MT_SHIFT[e * 8 + (dx | (dy << 1) | (dz << 2))];
            const int r2 = (cj - dy) * RX + (ci - dx);
            const int z2 = c - dz;
            const int cls2 = (gi - dx == 0) | ((gj - dy == 0) << 1) | ((gk - dz == 0) << 2);
            uint32_t oc;
            int64_t off;
            if (z2 >= 0) {
              oc = sm.ccase[r2 * kWave + z2];
              off = rbv[r2] + sm.cpv[r2 * kWave + z2];
            } else {
              oc = sm.pcase[r2];
              off = rbv[r2] - (int64_t)MT_NOWN[cls2 * 256 + oc];
            }
            id[m] = off + (int64_t)MT_RANK[(cls2 * 256 + oc) * 19 + e2] + 1;
          }
          if (fb + q < totf) {
            int64_t* out = faces + (fb + q) * 3;
            out[0] = id[0];
            out[1] = id[1];
            out[2] = id[2];
          }
        }
      } else {
        // SN vertex: mean of the crossing points over MC edges 1..12,
        // summed in promote(M, CX) per axis
        float fx = 0.f, fy = 0.f, fz = 0.f;
        double dx = 0.0, dy = 0.0, dz = 0.0;
        const bool sx = sizeof(M) == 4 && cx32, sy = sizeof(M) == 4 && cy32,
                   sz = sizeof(M) == 4 && cz32;
        int n = 0;
        for (int e = 0; e < 12; ++e) {
          const uint32_t ew = sm.lut_e[e];
          const int ca = ew & 0xF, cb = ew >> 4;
          if (((cs >> ca) & 1) == ((cs >> cb) & 1)) continue;
          ++n;
          const T va = pc[mc_dz(ca) * PS + mc_dy(ca) * SX + mc_dx(ca)];
          const T vbv = pc[mc_dz(cb) * PS + mc_dy(cb) * SX + mc_dx(cb)];
          const M mu = (iso - (M)va) / (M)(vbv - va);
          const double qx = lerp_axis<M>(cx32, mu, sm.cx[ci + mc_dx(ca)], sm.cx[ci + mc_dx(cb)]);
          const double qy = lerp_axis<M>(cy32, mu, sm.cy[cj + mc_dy(ca)], sm.cy[cj + mc_dy(cb)]);
          const double qz =
              lerp_axis<M>(cz32, mu, czs[c + H + mc_dz(ca)], czs[c + H + mc_dz(cb)]);
          if (sx) fx = fx + (float)qx; else dx = dx + qx;
          if (sy) fy = fy + (float)qy; else dy = dy + qy;
          if (sz) fz = fz + (float)qz; else dz = dz + qz;
        }
        if (vb < totv) {
          Out* out = verts + vb * 3;
          out[0] = (Out)(sx ? (double)(fx / (float)n) : dx / (double)n);
          out[1] = (Out)(sy ? (double)(fy / (float)n) : dy / (double)n);
          out[2] = (Out)(sz ? (double)(fz / (float)n) : dz / (double)n);
        }
        // SN faces: one quad per crossing axis edge at the first corner
        const uint32_t c0 = cs & 1;
        const int64_t me = vb + 1;
        int q = 0;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
          const uint32_t other = ax == 0 ? (cs >> 1) & 1 : ax == 1 ? (cs >> 3) & 1 : (cs >> 4) & 1;
          const int u = (ax + 1) % 3, w = (ax + 2) % 3;
          const int gidx[3] = {gi, gj, gk};
          if (c0 == other || gidx[u] < 1 || gidx[w] < 1) continue;
          auto cell_id = [&](int du, int dw) -> int64_t {
            int d[3] = {0, 0, 0};
            d[u] = du;
            d[w] = dw;
            const int r2 = (cj - d[1]) * RX + (ci - d[0]);
            const int z2 = c - d[2];
            if (z2 >= 0) return rbv[r2] + sm.cpv[r2 * kWave + z2] + 1;
            return rbv[r2];  // the active cell just below the chunk
          };
          const int64_t q1 = cell_id(1, 0), q2 = cell_id(1, 1), q3 = cell_id(0, 1);
          const int64_t a1 = c0 ? q1 : q3, a3 = c0 ? q3 : q1;
          if (fb + q < totf) {
            int64_t* out = faces + (fb + q) * 3;
            out[0] = me;
            out[1] = a1;
            out[2] = q2;
          }
          if (fb + q + 1 < totf) {
            int64_t* out = faces + (fb + q + 1) * 3;
            out[0] = me;
            out[1] = q2;
            out[2] = a3;
          }
          q += 2;
        }
      }
    };

    // Phase 2: classify (MC) or re-read (MT, SN) the rows this block emits,
    // scan each row, and collect its active cells; emit a full list with
    // one cell per lane.  Two rows are classified together so their LDS
    // and DPP latencies overlap.
    struct RowScan {
      uint32_t cs, packed, inc, tot;
      bool emit;
    };
    auto scan_row = [&](int r) {
      RowScan o{0u, 0u, 0u, 0u, false};
      const int ci = r % RX, cj = r / RX;
      const int gi = x0 + ci, gj = y0 + cj;
      if (ci < H || cj < H || gi >= g.nx - 1 || gj >= g.ny - 1) return o;
      o.emit = true;
      uint32_t cs, nv, nf;
      if constexpr (METHOD == MC) {
        const T* p = smp + cj * SX + ci;
        const uint32_t lo = face_code<MC, T, SX>(p + lane * PS, thr);
        const uint32_t top = face_code<MC, T, SX>(p + kWave * PS, thr);
        const uint32_t hi = next_lane(lo, top);
        cs = zvalid ? (lo | (hi << 4)) : 0u;
        nv = (uint32_t)(sm.lut_v[cs] & 0xF);
        nf = (uint32_t)(sm.lut_t[cs] & 0x7);
      } else if constexpr (METHOD == MT) {
        cs = sm.ccase[r * kWave + lane];
        const int cls = (gi == 0) | ((gj == 0) << 1) | ((kc == 0) << 2);
        nv = MT_NOWN[cls * 256 + cs];
        nf = MT_NF[cs];
      } else {
        cs = sm.ccase[r * kWave + lane];
        nv = (cs != 0 && cs != 255) ? 1u : 0u;
        nf = 0;
        if (nv) {
          const uint32_t c0 = cs & 1;
          nf += (c0 != ((cs >> 1) & 1) && gj >= 1 && kc >= 1) ? 2 : 0;
          nf += (c0 != ((cs >> 3) & 1) && kc >= 1 && gi >= 1) ? 2 : 0;
          nf += (c0 != ((cs >> 4) & 1) && gi >= 1 && gj >= 1) ? 2 : 0;
        }
      }
      o.cs = cs;
      o.packed = nv | (nf << 16);
      o.inc = wave_scan_dpp(o.packed);
      o.tot = (uint32_t)__builtin_amdgcn_readlane((int)o.inc, kWave - 1);
      return o;
    };
    int n = 0;  // wave-uniform list length
    auto take_row = [&](int r, const RowScan& o) {
      if (lane == 0) {
        if (!H) rbv_next[r] = rbv[r] + (o.tot & 0xFFFF);  // MT/SN: set in phase 1
        rbf_next[r] = rbf[r] + (o.tot >> 16);
      }
      if (o.tot == 0) return;
      const uint32_t ex = o.inc - o.packed;
      const bool active = o.packed != 0;
      const uint64_t bal = __ballot(active);
      const int cnt = __popcll(bal);
      if (n + cnt > CAP) {  // list full: emit it first
        wave_sync();
        for (int k = lane; k < n; k += kWave) emit_cell(list[k]);
        wave_sync();
        n = 0;
      }
      if (active)
        list[n + lanes_below(bal)] = (uint64_t)r | ((uint64_t)lane << 8) |
                                     ((uint64_t)o.cs << 14) | ((uint64_t)(ex & 0xFFFF) << 22) |
                                     ((uint64_t)(ex >> 16) << 32);
      n += cnt;
    };
    for (int r = wv; r < NR; r += 2 * kWaves) {
      const RowScan o0 = scan_row(r);
      const bool two = r + kWaves < NR;
      const RowScan o1 = two ? scan_row(r + kWaves) : RowScan{0u, 0u, 0u, 0u, false};
      take_row(r, o0);
      if (two) take_row(r + kWaves, o1);
    }
    // The next brick must have landed before the barrier that ends this
    // chunk; wait for it here, before this chunk's last stores are issued,
    // so the wait does not also cover them.
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
    for (int k = lane; k < n; k += kWave) emit_cell(list[k]);
    lds_barrier();
  }
}

template <int METHOD, typename T, typename M, typename A, typename Out, int TX, int TY>
__global__ __launch_bounds__(kBlock) void k_emit(VolArgs a) {
  emit_tile<METHOD, T, M, A, Out, TX, TY>(a, xcd_tile(blockIdx.x, a.ntiles));
}

template <int METHOD, typename T, typename M, typename A, typename Out, int TX, int TY>
__global__ __launch_bounds__(kBlock) void k_emit_b(const VolArgs* __restrict__ vols,
                                                   const int32_t* __restrict__ emap,
                                                   int64_t gbase) {
  const int64_t b = gbase + blockIdx.x;
  const VolArgs a = vols[emap[b]];
  emit_tile<METHOD, T, M, A, Out, TX, TY>(a, b - a.eblk0);
}

// ---------------------------------------------------------------------------
// Synthetic fields (no reference counterpart): column-major, x fastest.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * 0x1.0p-53; }

__device__ __forceinline__ double lin(int64_t i, int64_t n, double a, double b) {
  const int64_t d = n - 1 > 1 ? n - 1 : 1;
  const double t = (double)i / (double)d;
  return (1.0 - t) * a + t * b;
}

// Trilinear value noise on an L^3 lattice of seeded Uniform(-1,1) values
// spanning the unit cube.
__device__ double value_noise(int64_t i, int64_t j, int64_t k, int64_t nx, int64_t ny,
                              int64_t nz, int L, uint64_t seed) {
  const int64_t idx[3] = {i, j, k}, n[3] = {nx, ny, nz};
  int lo[3];
  double fr[3];
  for (int q = 0; q < 3; ++q) {
    const int64_t d = n[q] - 1 > 1 ? n[q] - 1 : 1;
    const double u = (double)(idx[q] * (L - 1)) / (double)d;
    int a = (int)floor(u);
    if (a > L - 2) a = L - 2;
    if (a < 0) a = 0;
    lo[q] = a;
    fr[q] = u - (double)a;
  }
  double acc = 0.0;
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
    const uint64_t li = (uint64_t)(lo[0] + dx) + (uint64_t)L * ((uint64_t)(lo[1] + dy) +
                                                               (uint64_t)L * (uint64_t)(lo[2] + dz));
    const double val = 2.0 * u01(splitmix64(seed ^ (li * 0xD1B54A32D192ED03ull))) - 1.0;
    const double w = (dx ? fr[0] : 1.0 - fr[0]) * (dy ? fr[1] : 1.0 - fr[1]) *
                     (dz ? fr[2] : 1.0 - fr[2]);
    acc = acc + w * val;
  }
  return acc;
}

template <typename T>
__global__ void k_generate(T* __restrict__ out, int kind, int64_t nx, int64_t ny, int64_t nz,
                           uint64_t seed, double p0, double p1) {
  const int64_t n = nx * ny * nz;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
    double v;
    if (kind == 0) {  // sphere
      const double x = lin(i, nx, -1, 1), y = lin(j, ny, -1, 1), z = lin(k, nz, -1, 1);
      v = sqrt(x * x + y * y + z * z) - p0;
    } else if (kind == 1) {  // torus
      const double x = lin(i, nx, -1, 1), y = lin(j, ny, -1, 1), z = lin(k, nz, -1, 1);
      const double qq = sqrt(x * x + y * y) - p0;
      v = sqrt(qq * qq + z * z) - p1;
    } else if (kind == 2) {  // gyroid on [-p0, p0]^3
      const double x = lin(i, nx, -p0, p0), y = lin(j, ny, -p0, p0), z = lin(k, nz, -p0, p0);
      v = cos(x) * sin(y) + cos(y) * sin(z) + cos(z) * sin(x);
    } else if (kind == 3) {  // value noise
      v = value_noise(i, j, k, nx, ny, nz, 32, seed) +
          0.25 * value_noise(i, j, k, nx, ny, nz, 128, seed ^ 0xA5A5A5A5A5A5A5A5ull);
    } else {  // 16 Gaussian blobs
      const double x = lin(i, nx, -1, 1), y = lin(j, ny, -1, 1), z = lin(k, nz, -1, 1);
      double acc = 0.0;
      for (int b = 0; b < 16; ++b) {
        const uint64_t h = seed * 16 + (uint64_t)b;
        const double cx = 1.2 * u01(splitmix64(4 * h + 0)) - 0.6;
        const double cy = 1.2 * u01(splitmix64(4 * h + 1)) - 0.6;
        const double cz = 1.2 * u01(splitmix64(4 * h + 2)) - 0.6;
        const double sg = 0.05 + 0.15 * u01(splitmix64(4 * h + 3));
        const double d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
        acc = acc + exp(-d2 / (2.0 * sg * sg));
      }
      v = 0.5 - acc;
    }
    out[q] = (T)v;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.

CoordArgs coord_args(const Plan& p, const Geom& g, double* coords) {
  CoordArgs c{};
  c.out = coords;
  c.nx = g.nx;
  c.ny = g.ny;
  c.nz = g.nz;
  for (int k = 0; k < 3; ++k) {
    c.a[k] = p.a[k];
    c.b[k] = p.b[k];
  }
  c.mask = (p.cx[0] == K_F32) | ((p.cx[1] == K_F32) << 1) | ((p.cx[2] == K_F32) << 2);
  return c;
}

hipError_t launch_coords(const Plan& p, const Geom& g, double* coords, hipStream_t st) {
  const int n = g.nx + g.ny + g.nz;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_coords, dim3((n + 255) / 256), dim3(256), 0, st, coord_args(p, g, coords));
  return hipGetLastError();
}

hipError_t launch_scan(const Geom& g, const Workspace& w, hipStream_t st, int64_t* host,
                       const CoordArgs* coords) {
  const int64_t nt = scan_tiles(g.units);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(kBlock), 0, st, w.cnt, g.units,
                     w.tilesum);
  if (nt <= kScanFuseTiles) {
    hipLaunchKernelGGL(k_scan_apply_sums, dim3((unsigned)nt), dim3(kBlock), 0, st, w.cnt,
                       g.units, (const uint64_t*)w.tilesum, w.voff, w.foff, w.totals, host,
                       coords ? *coords : CoordArgs{});
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kScanTilesBlock), 0, st,
                     (const uint64_t*)w.tilesum, nt, w.tileoff, w.totals, host,
                     coords ? *coords : CoordArgs{});
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nt), dim3(kBlock), 0, st, w.cnt, g.units,
                     (const int64_t*)w.tileoff, w.voff, w.foff);
  return hipGetLastError();
}

// Brick shape (cells along x, y) of the generic emit kernel, per element
// type: 31x1 for Float32, 15x1 for Float64 (of the shapes 0: 31x2, 1: 31x1,
// 2: 15x2, 3: 15x1).
static int generic_shape(const Plan& p) { return p.T == K_F32 ? 1 : 3; }

static const int kShapeTX[4] = {31, 31, 15, 15}, kShapeTY[4] = {2, 1, 2, 1};

// Variants: 0..31 = k_emit_mc4 (see mc4_variant); 32 + shape = k_emit;
// 64 + v = k_emit_halo, tetrahedra and surface nets (see halo_variant);
// 128 + v = k_emit_mc5 (see mc5_variant); 192 + v = k_emit_h5, tetrahedra
// and surface nets with lanes along x (see h5_variant).  The family comes
// from the plan only (msh_set_kernel_family / msh_autotune).
int emit_variant(const Plan& p, VolArgs* a) {
  const int fam = p.family < 0 ? 2 : p.family;
  if (fam >= 2) {
    const int v = p.method == MC ? mc5_variant(p, a) : h5_variant(p, a);
    if (v >= 0) return (p.method == MC ? 128 : 192) + v;
  }
  if (fam >= 1) {
    const int v = p.method == MC ? mc4_variant(p, a) : halo_variant(p, a);
    if (v >= 0) return (p.method == MC ? 0 : 64) + v;
  }
  const int sh = generic_shape(p);
  a->nti = (a->g.nx - 1 + kShapeTX[sh] - 1) / kShapeTX[sh];
  a->ntj = (a->g.ny - 1 + kShapeTY[sh] - 1) / kShapeTY[sh];
  a->ntiles = (a->g.nx > 1 && a->g.ny > 1) ? (int64_t)a->nti * a->ntj * a->g.nzseg : 0;
  return 32 + sh;
}

int variant_family(int variant) {
  if (variant >= 128 && ((variant & 63) & 8)) return 3;  // half-depth bricks (MC, SN)
  if (variant >= 128) return 2;
  if (variant >= 64 || variant < 32) return 1;
  return 0;
}

void count_tiling(const Plan& p, VolArgs* a, bool split) { count4_tiling(p, a, split); }

VolArgs make_vol(const Plan& p, const Geom& g, const void* field, const Workspace& w) {
  VolArgs a{};
  a.f = field;
  a.g = g;
  a.thr = p.thr;
  a.iso = p.isoM;
  a.eps = p.epsA;
  a.ome = p.omeA;
  a.cxmask = (p.cx[0] == K_F32) | ((p.cx[1] == K_F32) << 1) | ((p.cx[2] == K_F32) << 2);
  a.method = p.method;
  for (int k = 0; k < 3; ++k) {
    a.ca[k] = p.a[k];
    a.cb[k] = p.b[k];
  }
  a.coords = w.coords;
  a.cnt = w.cnt;
  a.zmask = w.zmask;
  a.voff = w.voff;
  a.foff = w.foff;
  a.sz = w.totals;
  a.ubase = 0;
  count_tiling(p, &a, true);
  return a;
}

hipError_t launch_count(const Plan& p, const VolArgs& a, hipStream_t st) {
  return launch_count4(p, a, st);
}

template <int METHOD, typename T, typename M, typename A, typename Out, int TX, int TY>
static hipError_t emit_t(const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                         int64_t gbase, int64_t nblocks, hipStream_t st) {
  if (vols)
    hipLaunchKernelGGL((k_emit_b<METHOD, T, M, A, Out, TX, TY>), dim3((unsigned)nblocks),
                       dim3(kBlock), 0, st, vols, emap, gbase);
  else
    hipLaunchKernelGGL((k_emit<METHOD, T, M, A, Out, TX, TY>), dim3((unsigned)a.ntiles),
                       dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

template <int METHOD, typename T, typename M, typename A, typename Out>
static hipError_t emit_shape(int sh, const VolArgs& a, const VolArgs* vols, const int32_t* emap,
                             int64_t gbase, int64_t nblocks, hipStream_t st) {
  if constexpr (METHOD == MC) {
    switch (sh) {
      case 0: return emit_t<MC, T, M, A, Out, 31, 2>(a, vols, emap, gbase, nblocks, st);
      case 1: return emit_t<MC, T, M, A, Out, 31, 1>(a, vols, emap, gbase, nblocks, st);
      case 2: return emit_t<MC, T, M, A, Out, 15, 2>(a, vols, emap, gbase, nblocks, st);
      default: return emit_t<MC, T, M, A, Out, 15, 1>(a, vols, emap, gbase, nblocks, st);
    }
  } else if constexpr (sizeof(T) == 4) {
    return emit_t<METHOD, T, M, A, Out, 31, 1>(a, vols, emap, gbase, nblocks, st);
  } else {
    return emit_t<METHOD, T, M, A, Out, 15, 1>(a, vols, emap, gbase, nblocks, st);
  }
}

template <int METHOD>
static hipError_t emit_types(const Plan& p, int sh, const VolArgs& a, const VolArgs* vols,
                             const int32_t* emap, int64_t gbase, int64_t nblocks,
                             hipStream_t st) {
#define MSH_EMIT_ARGS sh, a, vols, emap, gbase, nblocks, st
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32 && p.Out == K_F32)
    return emit_shape<METHOD, float, float, float, float>(MSH_EMIT_ARGS);
  if (p.T == K_F32 && p.M == K_F32 && p.A == K_F32)
    return emit_shape<METHOD, float, float, float, double>(MSH_EMIT_ARGS);
  if (p.T == K_F32 && p.M == K_F32)
    return emit_shape<METHOD, float, float, double, double>(MSH_EMIT_ARGS);
  if (p.T == K_F32) return emit_shape<METHOD, float, double, double, double>(MSH_EMIT_ARGS);
  return emit_shape<METHOD, double, double, double, double>(MSH_EMIT_ARGS);
#undef MSH_EMIT_ARGS
}

static hipError_t emit_any(const Plan& p, int variant, const VolArgs& a, const VolArgs* vols,
                           const int32_t* emap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                           hipStream_t st) {
  if (variant < 32) {
    if (vols) return launch_emit_mc4_batch(p, variant, vols, emap, gbase, nblocks, st);
    return launch_emit_mc4(p, variant, a, st);
  }
  if (variant >= 192) {
    if (vols) return launch_emit_h5_batch(p, variant - 192, vols, emap, gbase, nblocks, ctr, st);
    return launch_emit_h5(p, variant - 192, a, st);
  }
  if (variant >= 128) {
    if (vols) return launch_emit_mc5_batch(p, variant - 128, vols, emap, gbase, nblocks, ctr, st);
    return launch_emit_mc5(p, variant - 128, a, st);
  }
  if (variant >= 64) {
    if (vols) return launch_emit_halo_batch(p, variant - 64, vols, emap, gbase, nblocks, st);
    return launch_emit_halo(p, variant - 64, a, st);
  }
  if (!vols && a.ntiles == 0) return hipSuccess;
  if (vols && nblocks == 0) return hipSuccess;
  const int sh = variant - 32;
  Plan q = p;
  if (p.method != MT) q.A = q.M;  // eps is used by Marching Tetrahedra only
  if (p.method == MC) return emit_types<MC>(q, sh, a, vols, emap, gbase, nblocks, st);
  if (p.method == SN) return emit_types<SN>(q, sh, a, vols, emap, gbase, nblocks, st);
  return emit_types<MT>(q, sh, a, vols, emap, gbase, nblocks, st);
}

hipError_t launch_emit(const Plan& p, const VolArgs& a0, hipStream_t st) {
  VolArgs a = a0;
  const int variant = emit_variant(p, &a);
  return emit_any(p, variant, a, nullptr, nullptr, 0, 0, nullptr, st);
}

hipError_t launch_coords_batch(const VolArgs* vols, int n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_coords_b, dim3((unsigned)n), dim3(256), 0, st, vols);
  return hipGetLastError();
}

hipError_t launch_fill_maps(const VolArgs* vols, int n, int32_t* cmap, int32_t* emap,
                            hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_maps, dim3((unsigned)n), dim3(256), 0, st, vols, cmap, emap);
  return hipGetLastError();
}

hipError_t launch_count_batch(const Plan& p, const VolArgs* vols, const int32_t* cmap,
                              int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  return launch_count4_batch(p, vols, cmap, gbase, nblocks, ctr, st);
}

hipError_t launch_zero_u32(uint32_t* p, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

unsigned persistent_grid(const void* kernel, int64_t ntiles) {
  static std::mutex mu;
  static std::map<std::pair<int, const void*>, int64_t> cache;  // (device, kernel) -> blocks
  int dev = 0;
  (void)hipGetDevice(&dev);
  int64_t resident = 0;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find({dev, kernel});
    if (it != cache.end()) {
      resident = it->second;
    } else {
      int per_cu = 0, cus = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess)
        per_cu = 1;
      if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
      resident = (int64_t)(per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
      cache[{dev, kernel}] = resident;
    }
  }
  return (unsigned)(ntiles < resident ? ntiles : resident);
}

hipError_t launch_sizes_batch(const VolArgs* vols, int n, const Workspace& w, int64_t units,
                              int64_t* sizes, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sizes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, vols, n,
                     w.voff, w.foff, w.totals, units, sizes);
  return hipGetLastError();
}

hipError_t launch_emit_batch(const Plan& p, int variant, const VolArgs* vols,
                             const int32_t* emap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                             hipStream_t st) {
  return emit_any(p, variant, VolArgs{}, vols, emap, gbase, nblocks, ctr, st);
}

hipError_t launch_generate(int kind, int dtype, int64_t nx, int64_t ny, int64_t nz,
                           uint64_t seed, double p0, double p1, void* out, hipStream_t st) {
  const int64_t n = nx * ny * nz;
  if (n == 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (dtype == K_F32)
    hipLaunchKernelGGL(k_generate<float>, dim3((unsigned)blocks), dim3(256), 0, st, (float*)out,
                       kind, nx, ny, nz, seed, p0, p1);
  else
    hipLaunchKernelGGL(k_generate<double>, dim3((unsigned)blocks), dim3(256), 0, st,
                       (double*)out, kind, nx, ny, nz, seed, p0, p1);
  return hipGetLastError();
}

}  // namespace msh
