// lean.hip -- the count kernel of all methods and the Marching Cubes emit
// kernel, written to keep vector-ALU work per cell small (these passes are
// bound by VALU issue as much as by HBM).
//
// Both classify cells from "below iso" bit masks instead of comparing each
// cell's 8 corners: a sample is compared once, and the 4 bits of a cell
// face are pulled out of a mask with one shift.  Bits are kept in
// "natural" order inside each nibble -- (x,y) (x+1,y) (x,y+1) (x+1,y+1) --
// and the lookup tables are permuted once per block to that order.
#include <stdlib.h>

#include <type_traits>

#include "device.h"

#define MSH_LUT static __constant__
#include "luts.h"

namespace msh {
namespace {

// Cells per emit tile along x: 60 keeps every tile's first sample 16-byte
// aligned (when the row pitch is) while 61 samples fit a 64-lane row.
constexpr int kEmitTX = 60;

template <typename T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <>
__device__ __forceinline__ float buffer_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
template <>
__device__ __forceinline__ double buffer_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// ---------------------------------------------------------------------------
// Count.  The 4 waves of a block sit side by side along x; wave w covers
// sample columns [xb, xb+64) (cells xb..xb+62) and TY cell rows from yb, and
// marches the z segment of its tile.  On a volume too small to give every
// SIMD a few such waves, 2 or 4 waves share one x span instead (a.czl), each
// marching its own run of whole z blocks, and their sums meet in LDS: a wave
// is latency-bound (one group of planes in flight behind the other), so
// shorter runs on more waves finish sooner.  Per plane a lane loads TY+1 samples
// (buffer loads: the address is a per-row offset fixed for the whole run
// plus a per-plane base in scalar registers), packs their below-iso bits at
// even positions, and ORs in the x+1 column's bits from the next lane at
// odd positions; cell row r's face nibble is then bits [2r, 2r+4).
template <typename T, int METHOD, int TY, int ZU>
__device__ __forceinline__ void count4_body(const VolArgs& a, int64_t t,
                                            const uint32_t* __restrict__ lut);

// The per-cell count tables of the three methods (vertices | faces << 16,
// by class << 8 | natural case; Marching Cubes has no classes), built once
// per device (tables_setup in capi.cpp, at context creation); each count block
// copies its method's table to LDS.
static __device__ __align__(16) uint32_t g_count_lut[3][2048];

__global__ __launch_bounds__(kBlock) void k_count_tables() {
  for (int i = threadIdx.x; i < 3 * 2048; i += kBlock) {
    const int method = i / 2048, c = i % 2048;
    const uint32_t cs = c & 255, cls = c >> 8;
    uint32_t v;
    if (method == MC) {
      const uint32_t m = nat_to_mc(cs);
      v = (uint32_t)(MC_CASE_VERTS[m] & 0xF) | ((uint32_t)(MC_CASE_TRIS[m] & 0x7) << 16);
    } else if (method == MT) {
      // cls: first x layer | first y layer << 1 | first z layer << 2
      const uint32_t m = nat_to_mt(cs);
      v = (uint32_t)MT_NOWN[cls * 256 + m] | ((uint32_t)MT_NF[m] << 16);
    } else {
      // cls: x >= 1 | y >= 1 << 1 | z >= 1 << 2 (a quad needs the cells at -1)
      const uint32_t m = nat_to_mc(cs);
      const uint32_t nv = (m != 0 && m != 255) ? 1u : 0u;
      uint32_t nf = 0;
      if (nv) {
        const uint32_t c0 = m & 1;
        nf += (c0 != ((m >> 1) & 1) && (cls & 2) && (cls & 4)) ? 2 : 0;  // x edge
        nf += (c0 != ((m >> 3) & 1) && (cls & 4) && (cls & 1)) ? 2 : 0;  // y edge
        nf += (c0 != ((m >> 4) & 1) && (cls & 1) && (cls & 2)) ? 2 : 0;  // z edge
      }
      v = nv | (nf << 16);
    }
    g_count_lut[method][c] = v;
  }
}

// The count kernels' tile loop: the per-cell count table goes to LDS once
// per block, then the tiles of `src` are counted one after another.
template <typename T, int METHOD, int TY, int ZU, class Src>
__device__ __forceinline__ void count4_run(Src src) {
  constexpr int NL = METHOD == MC ? 256 : 2048;
  __shared__ __align__(16) uint32_t lut[NL];  // vertices | faces << 16, by class << 8 | natural case
  {
    const uint4* __restrict__ from = (const uint4*)g_count_lut[METHOD];
    for (int q = threadIdx.x; q < NL / 4; q += kBlock) ((uint4*)lut)[q] = from[q];
  }
  __syncthreads();
  if constexpr (Src::kOne) {
    count4_body<T, METHOD, TY, ZU>(*src.p, src.t, lut);
  } else {
    VolArgs a;
    int64_t t;
    while (src(a, t)) count4_body<T, METHOD, TY, ZU>(a, t, lut);
  }
}

template <typename T, int METHOD, int TY, int ZU>
__device__ __forceinline__ void count4_body(const VolArgs& a, int64_t t,
                                            const uint32_t* __restrict__ lut) {
  static_assert(2 * (TY + 1) <= 32, "mask must fit 32 bits");
  const T* __restrict__ f = (const T*)a.f;
  const Geom g = a.g;
  const T thr = (T)a.thr;
  uint32_t* __restrict__ cnt = a.cnt + a.ubase;
  const int nbx = a.cnbx, nby = a.cnby;
  const int lane = lane_id(), wv = wave_id();
  const int by = (int)(t % nby);
  const int64_t rest = t / nby;
  const int bx = (int)(rest % nbx);
  const int s = (int)(rest / nbx);
  const int czl = a.czl;  // uniform across the block
  const int wx = wv >> czl, wz = wv & ((1 << czl) - 1);
  const int xb = (bx * (kWaves >> czl) + wx) * (kWave - 1);
  const int yb = by * TY;
  const bool xin = xb < g.nx - 1;
  if (czl == 0 && !xin) return;
  const int zs0 = s * g.zseg;  // the segment's start: mask bits count from it
  const int zend = min(zs0 + g.zseg, g.nz - 1);
  const int zsh = zmask_shift(g.zseg);
  // this wave's cells [zs, ze): whole z blocks of the segment (a wave
  // beyond the volume's x end takes none, but meets the block's barrier)
  int zs = zs0, ze = zend;
  if (czl) {
    const int nblk = (zend - zs0 + (1 << zsh) - 1) >> zsh;
    const int b0 = (wz * nblk) >> czl, b1 = ((wz + 1) * nblk) >> czl;
    zs = min(zend, zs0 + (b0 << zsh));
    ze = xin ? min(zend, zs0 + (b1 << zsh)) : zs;
  }
  const bool xvalid = lane < kWave - 1 && xb + lane < g.nx - 1;
  const int xs = min(xb + lane, g.nx - 1);
  const int64_t pstride = (int64_t)g.nx * g.ny;
  const uint32_t pbytes = (uint32_t)(pstride * (int64_t)sizeof(T));

  uint32_t voff[TY + 1];
#pragma unroll
  for (int r = 0; r <= TY; ++r)
    voff[r] = ((uint32_t)min(yb + r, g.ny - 1) * (uint32_t)g.nx + (uint32_t)xs) * sizeof(T);
  // class bits of the per-cell table (MT, SN): x from the lane, y per row
  // (the same in every lane, so they stay in scalar registers), z per plane
  uint32_t clsx = 0, clsy[TY];
  if (METHOD == MT) clsx = (xb + lane == 0) ? 1u : 0u;
  if (METHOD == SN) clsx = (xb + lane >= 1) ? 1u : 0u;
#pragma unroll
  for (int r = 0; r < TY; ++r)
    clsy[r] = METHOD == MT ? ((yb + r == 0) ? 2u : 0u) : ((yb + r >= 1) ? 2u : 0u);

  auto plane_bits = [&](const T* v) {
    uint32_t w = 0;
#pragma unroll
    for (int r = 0; r <= TY; ++r) w |= (uint32_t)(v[r] < thr) << (2 * r);
    return w | (next_lane(w, 0u) << 1);
  };

  uint32_t acc[TY];
#pragma unroll
  for (int r = 0; r < TY; ++r) acc[r] = 0;
  // z blocks where some cell of this lane's TY rows has outputs: a block is
  // marked when the rows' running counts grew across it (the packed sum
  // of a block's growth stays far below 2^32, so it cannot wrap to zero)
  uint64_t zm = 0;
  uint32_t tprev = 0;

  auto load_group = [&](int k, T (&v)[ZU][TY + 1]) {
#pragma unroll
    for (int u = 0; u < ZU; ++u) {
      const __amdgpu_buffer_rsrc_t rs =
          buffer_rsrc(f + (int64_t)min(k + 1 + u, g.nz - 1) * pstride, pbytes);
#pragma unroll
      for (int r = 0; r <= TY; ++r) v[u][r] = buffer_load<T>(rs, voff[r]);
    }
  };
  uint32_t w0;
  auto use_group = [&](int k, const T (&v)[ZU][TY + 1]) {
#pragma unroll
    for (int u = 0; u < ZU; ++u) {
      if (k + u >= ze) break;
      const uint32_t w1 = plane_bits(v[u]);
      const int kc = k + u;
      const uint32_t clxz = clsx | (METHOD == MT ? ((kc == 0) ? 4u : 0u) : ((kc >= 1) ? 4u : 0u));
#pragma unroll
      for (int r = 0; r < TY; ++r) {
        const uint32_t cs = ((w0 >> (2 * r)) & 0xF) | (((w1 >> (2 * r)) & 0xF) << 4);
        const uint32_t idx = METHOD == MC ? cs : (cs | ((clsy[r] | clxz) << 8));
        acc[r] += lut[idx];
      }
      w0 = w1;
    }
  };
  {
    T v0[TY + 1];
    const __amdgpu_buffer_rsrc_t rs = buffer_rsrc(f + (int64_t)zs * pstride, pbytes);
#pragma unroll
    for (int r = 0; r <= TY; ++r) v0[r] = buffer_load<T>(rs, voff[r]);
    w0 = plane_bits(v0);
  }
  // after the group ending before cell k: close the z block when it ends
  // there (blocks are whole groups; the last one may be short)
  auto close_block = [&](int k) {
    if (((k - zs0) & ((1 << zsh) - 1)) == 0 || k >= ze) {
      uint32_t t = 0;
#pragma unroll
      for (int r = 0; r < TY; ++r) t += acc[r];
      zm |= (uint64_t)(t != tprev) << ((k - 1 - zs0) >> zsh);
      tprev = t;
    }
  };
  // two register sets: the next group's loads are in flight while the
  // current group is classified
  T va[ZU][TY + 1], vb[ZU][TY + 1];
  int k = zs;
  if (k < ze) load_group(k, va);
  while (k < ze) {
    if (k + ZU < ze) load_group(k + ZU, vb);
    __builtin_amdgcn_sched_barrier(0);
    use_group(k, va);
    k += ZU;
    close_block(k);
    if (k >= ze) break;
    if (k + ZU < ze) load_group(k + ZU, va);
    __builtin_amdgcn_sched_barrier(0);
    use_group(k, vb);
    k += ZU;
    close_block(k);
  }
  if (czl) {
    // the waves of one x span add their rows' counts and OR their masks
    // (disjoint z blocks); the wave with the first run stores them.  The
    // next tile of a persistent block writes here only after the tile
    // queue's barrier, which every wave reaches after its reads below.
    __shared__ uint32_t pacc[kWaves][TY][kWave];
    __shared__ uint64_t pzm[kWaves][kWave];
    if (wz) {
#pragma unroll
      for (int r = 0; r < TY; ++r) pacc[wv][r][lane] = acc[r];
      pzm[wv][lane] = zm;
    }
    lds_barrier();
    if (wz) return;
    for (int q = 1; q < (1 << czl); ++q) {
#pragma unroll
      for (int r = 0; r < TY; ++r) acc[r] += pacc[wv + q][r][lane];
      zm |= pzm[wv + q][lane];
    }
    if (!xin) return;
  }
  if (!xvalid) return;
#pragma unroll
  for (int r = 0; r < TY; ++r) {
    const int j = yb + r;
    if (j < g.ny - 1) cnt[((int64_t)(xb + lane) * (g.ny - 1) + j) * g.nzseg + s] = acc[r];
  }
  // one mask for the lane's TY rows (zmask_at)
  a.zmask[a.ubase + ((int64_t)(xb + lane) * nby + by) * g.nzseg + s] = zm;
}

template <typename T, int METHOD, int TY, int ZU>
__global__ __launch_bounds__(kBlock) void k_count4(VolArgs a) {
  count4_run<T, METHOD, TY, ZU>(OneTile{&a, xcd_tile(blockIdx.x, a.cntiles), false});
}

// Batched, persistent: the launch's blocks take count tiles (block gbase + i
// of the batch; the map names its volume, whose tiles start at block cblk0)
// from a work queue.
template <typename T, int METHOD, int TY, int ZU>
__global__ __launch_bounds__(kBlock) void k_count4_b(const VolArgs* __restrict__ vols,
                                                     const int32_t* __restrict__ cmap,
                                                     int64_t gbase, int64_t ntiles,
                                                     uint32_t* __restrict__ ctr) {
  count4_run<T, METHOD, TY, ZU>(TileQueue<false>{vols, cmap, gbase, ntiles, ctr, 0});
}

// ---------------------------------------------------------------------------
// Emit, Marching Cubes.
//
// A block owns 63 cell rows along x at one y (samples x in [ib, ib+64),
// y in {jb, jb+1}) and walks its z segment in chunks of ZC cells.  Per
// chunk:
//  * the brick (ZC+1 planes x 2 rows x 64 samples) arrives by LDS DMA, one
//    buffer_load...lds per sample row (addresses: a scalar base per row plus
//    lane*4), double-buffered so the next chunk is in flight meanwhile;
//  * one ballot per sample row turns it into a 64-bit below-iso mask;
//  * each wave step takes R = 64/ZC rows: lane (h, z) is cell (ci = R*step
//    + h, jb, zc + z).  The case comes from 4 mask reads, the row's output
//    slots from a segmented DPP scan (the reference's emission order inside
//    a row), and active cells are appended to the wave's list with their
//    slot prefix;
//  * the list is emitted slot-parallel: lane o takes vertex slot o (and face
//    slot o), finds its cell by binary search over the list prefixes, and
//    writes output rbv[row] + slot -- consecutive lanes write consecutive
//    outputs of a row.
template <typename T, typename M, typename Out, int ZC, int CY, bool VEC, int CXM>
__device__ __forceinline__ void emit_mc4_tile(const VolArgs& a, int64_t t) {
  constexpr int TX = kEmitTX, SX = 64, SY = CY + 1;
  constexpr int R = kWave / ZC;
  constexpr int NP = ZC + 1;
  constexpr int DW = (int)(sizeof(T) / 4);
  constexpr int RD = SX * DW;  // dwords per sample row
  constexpr int NRW = TX * CY;  // cell rows (x, y) of the tile
  static_assert(TX % R == 0, "a wave step must not straddle two y rows");
  constexpr int NSTEP = NRW / R;
  constexpr int CAP = 128;
  constexpr int VW = CAP * 12 / 64, FW = CAP * 5 / 64;  // slot bitmap words
  struct Smem {
    uint32_t smp[2][NP * SY * RD];
    double cz[2][NP];
    double cx[SX], cy[SY];
    uint64_t mask[NP * SY];  // [plane][y], bit x
    uint32_t cnt[256];
    uint64_t lv[256], lt[256];
    uint32_t le[16];  // per edge: brick offsets of its corners a, b and their (dx,dy,dz)
    int64_t rbv[2][NRW], rbf[2][NRW];  // output offset of each row at this / next chunk
    uint64_t list[kWaves][CAP];
    uint16_t lpv[kWaves][CAP], lpf[kWaves][CAP];
    uint64_t vbits[kWaves][VW], fbits[kWaves][FW];  // bit o: slot o starts an entry
  };
  __shared__ Smem sm;
  const T* __restrict__ f = (const T*)a.f;
  Out* __restrict__ verts = (Out*)a.verts;
  int64_t* __restrict__ faces = a.faces;
  const Geom g = a.g;
  const int lane = lane_id(), wv = wave_id();
  const int jb = (int)(t % a.ntj) * CY;
  const int64_t rest = t / a.ntj;
  const int ib = (int)(rest % a.nti) * TX;
  const int s = (int)(rest / a.nti);
  const int zs = s * g.zseg, ze = min(zs + g.zseg, g.nz - 1);
  const T thr = (T)a.thr;
  const M iso = (M)a.iso;
  // axis element types: compile-time when CXM >= 0 (all Float64: 0, all
  // Float32: 7), else read from the arguments
  const int cxm = CXM >= 0 ? CXM : a.cxmask;
  const bool cx32 = cxm & 1, cy32 = (cxm >> 1) & 1, cz32 = (cxm >> 2) & 1;
  const double* zp = a.coords + g.nx + g.ny;
  const int64_t vbase = a.sz[0], totv = a.sz[1], fbase = a.sz[2], totf = a.sz[3];
  const int64_t nelem = (int64_t)g.nx * g.ny * g.nz;
  const int64_t pstride = (int64_t)g.nx * g.ny;

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
  bool emits = false;
  for (int rr = threadIdx.x; rr < NRW; rr += kBlock) {
    const int ci = rr % TX, cj = rr / TX;
    int64_t bv = 0, bf = 0;
    if (ib + ci < g.nx - 1 && jb + cj < g.ny - 1) {
      const int64_t u = a.ubase + ((int64_t)(ib + ci) * (g.ny - 1) + jb + cj) * g.nzseg + s;
      bv = a.voff[u] - vbase;
      bf = a.foff[u] - fbase;
      emits = emits || a.cnt[u] != 0;
    }
    sm.rbv[0][rr] = bv;
    sm.rbf[0][rr] = bf;
  }
  // a tile whose rows the count found empty writes nothing: skip it
  // before fetching any of its bricks (sparse fields have many)
  if (!__syncthreads_or(emits)) return;
  for (int c = threadIdx.x; c < SX; c += kBlock) sm.cx[c] = a.coords[min(ib + c, g.nx - 1)];
  if (threadIdx.x < SY) sm.cy[threadIdx.x] = a.coords[g.nx + min(jb + (int)threadIdx.x, g.ny - 1)];

  // DMA layout.  VEC: 16 bytes per lane, so one instruction fills 1 KiB of
  // LDS = 1024/(4*RD) whole sample rows; the per-lane byte offsets relative
  // to the chunk's first sample are the same for every chunk and are
  // computed once.  Otherwise one dword per lane, one instruction per
  // 64-dword piece of a row.
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

  uint64_t* list = sm.list[wv];
  uint16_t* lpv = sm.lpv[wv];
  uint16_t* lpf = sm.lpf[wv];
  uint64_t* vbits = sm.vbits[wv];
  uint64_t* fbits = sm.fbits[wv];
  for (int q = lane; q < VW; q += kWave) vbits[q] = 0;
  for (int q = lane; q < FW; q += kWave) fbits[q] = 0;
  const int h = lane / ZC, z = lane % ZC;

  if (zs < ze) fetch(zs, 0);
  __builtin_amdgcn_s_waitcnt(0);
  lds_barrier();
  int buf = 0;
  for (int zc = zs; zc < ze; zc += ZC, buf ^= 1) {
    if (zc + ZC < ze) fetch(zc + ZC, buf ^ 1);
    const uint32_t* smp = sm.smp[buf];
    const double* czs = sm.cz[buf];
    const int64_t* rbv = sm.rbv[buf];
    const int64_t* rbf = sm.rbf[buf];
    int64_t* rbv_next = sm.rbv[buf ^ 1];
    int64_t* rbf_next = sm.rbf[buf ^ 1];

    // below-iso masks of this chunk's sample rows
    for (int row = wv; row < NP * SY; row += kWaves) {
      const T v = *(const T*)(smp + row * RD + lane * DW);
      const uint64_t b = __ballot(v < thr);
      if (lane == 0) sm.mask[row] = b;
    }
    lds_barrier();

    int n = 0;
    uint32_t V = 0, F = 0;
    auto flush = [&]() {
      // slot o belongs to the last entry starting at or before o: count
      // the start bits up to o
      int before = 0;
      for (uint32_t o0 = 0; o0 < V; o0 += kWave) {
        const uint32_t o = o0 + lane;
        const uint64_t word = vbits[o0 >> 6];
        const uint64_t upto = word & (~0ull >> (kWave - 1 - lane));
        const int k = before + __popcll(upto) - 1;
        before += __popcll(word);
        if (o >= V) break;
        const uint64_t e = list[k];
        const int ci = (int)(e & 63), cj = (int)((e >> 6) & 3), c = (int)((e >> 8) & 31);
        const uint32_t cs = (uint32_t)((e >> 13) & 0xFF), pv = (uint32_t)((e >> 21) & 0x1FF);
        const uint32_t q = o - lpv[k];
        const uint32_t ed = (uint32_t)(sm.lv[cs] >> (4 + 4 * q)) & 0xF;
        const uint32_t ew = sm.le[ed];
        const uint32_t* cell = smp + c * (SY * RD) + cj * RD + ci * DW;
        const T va = *(const T*)(cell + (ew & 0x3FF)), vb = *(const T*)(cell + ((ew >> 10) & 0x3FF));
        const M mu = (iso - (M)va) / (M)(vb - va);
        const double px = lerp_axis<M>(cx32, mu, sm.cx[ci + ((ew >> 20) & 1)], sm.cx[ci + ((ew >> 23) & 1)]);
        const double py = lerp_axis<M>(cy32, mu, sm.cy[cj + ((ew >> 21) & 1)], sm.cy[cj + ((ew >> 24) & 1)]);
        const double pz = lerp_axis<M>(cz32, mu, czs[c + ((ew >> 22) & 1)], czs[c + ((ew >> 25) & 1)]);
        const int64_t vi = rbv[cj * TX + ci] + pv + q;
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
        const uint64_t e = list[k];
        const int rr = (int)((e >> 6) & 3) * TX + (int)(e & 63);
        const uint32_t cs = (uint32_t)((e >> 13) & 0xFF), pv = (uint32_t)((e >> 21) & 0x1FF),
                       pf = (uint32_t)((e >> 30) & 0x1FF);
        const uint32_t q = o - lpf[k];
        const uint64_t w = sm.lt[cs] >> (3 + 12 * q);
        const int64_t vb = rbv[rr] + pv + 1;  // 1-based
        const int64_t fi = rbf[rr] + pf + q;
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

    const int kc = zc + z;
    const bool zvalid = kc < ze;
    for (int st = wv; st < NSTEP; st += kWaves) {
      const int r0 = st * R, cj = r0 / TX;  // the step's y row (wave-uniform)
      const int rr = r0 + h, ci = r0 - cj * TX + h;
      const bool valid = zvalid && ib + ci < g.nx - 1 && jb + cj < g.ny - 1;
      const uint64_t* mz = sm.mask + z * SY + cj;
      const uint32_t lo = (uint32_t)((mz[0] >> ci) & 3) | ((uint32_t)((mz[1] >> ci) & 3) << 2);
      const uint32_t hi =
          (uint32_t)((mz[SY] >> ci) & 3) | ((uint32_t)((mz[SY + 1] >> ci) & 3) << 2);
      const uint32_t cs = valid ? (lo | (hi << 4)) : 0u;
      const uint32_t packed = sm.cnt[cs];
      const uint32_t inc = seg_scan_dpp<ZC>(packed);
      uint32_t tots[R];
#pragma unroll
      for (int r = 0; r < R; ++r)
        tots[r] = (uint32_t)__builtin_amdgcn_readlane((int)inc, r * ZC + ZC - 1);
      uint32_t before = 0, mine = tots[0], sum = 0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (r < h) before += tots[r];
        if (r == h) mine = tots[r];
        sum += tots[r];
      }
      if (z == 0) {
        rbv_next[rr] = rbv[rr] + (mine & 0xFFFF);
        rbf_next[rr] = rbf[rr] + (mine >> 16);
      }
      if (sum == 0) continue;
      const uint32_t ex_row = inc - packed;   // prefix inside the cell's row
      const uint32_t ex = ex_row + before;    // prefix inside the step
      const bool active = packed != 0;
      const uint64_t bal = __ballot(active);
      const int na = __popcll(bal);
      if (n + na > CAP) {
        wave_sync();
        flush();
        wave_sync();
        n = 0;
        V = 0;
        F = 0;
      }
      if (active) {
        const int k = n + (int)lanes_below(bal);
        list[k] = (uint64_t)(ci | (cj << 6)) | ((uint64_t)z << 8) | ((uint64_t)cs << 13) |
                  ((uint64_t)(ex_row & 0xFFFF) << 21) | ((uint64_t)(ex_row >> 16) << 30);
        const uint32_t sv = V + (ex & 0xFFFF), sf = F + (ex >> 16);
        lpv[k] = (uint16_t)sv;
        lpf[k] = (uint16_t)sf;
        atomicOr((unsigned long long*)&vbits[sv >> 6], 1ull << (sv & 63));
        atomicOr((unsigned long long*)&fbits[sf >> 6], 1ull << (sf & 63));
      }
      n += na;
      V += sum & 0xFFFF;
      F += sum >> 16;
    }
    // the next brick must land before the barrier that ends this chunk; wait
    // for it before this chunk's last stores are issued
    __builtin_amdgcn_s_waitcnt(0);
    wave_sync();
    flush();
    lds_barrier();
  }
}

template <typename T, typename M, typename Out, int ZC, int CY, bool VEC, int CXM>
__global__ __launch_bounds__(kBlock) void k_emit_mc4(VolArgs a) {
  emit_mc4_tile<T, M, Out, ZC, CY, VEC, CXM>(a, xcd_tile(blockIdx.x, a.ntiles));
}

template <typename T, typename M, typename Out, int ZC, int CY, bool VEC, int CXM>
__global__ __launch_bounds__(kBlock) void k_emit_mc4_b(const VolArgs* __restrict__ vols,
                                                       const int32_t* __restrict__ emap,
                                                       int64_t gbase) {
  const int64_t b = gbase + blockIdx.x;
  const VolArgs a = vols[emap[b]];
  emit_mc4_tile<T, M, Out, ZC, CY, VEC, CXM>(a, b - a.eblk0);
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side.

// planes per load group (two groups in flight): 2 for Float32, 4 for
// Float64 measured best at 1024^3 (Float32 1.14 vs 1.18 ms with 4)
template <typename T>
static constexpr int count_zu() { return sizeof(T) == 4 ? 2 : 4; }

hipError_t launch_count_tables(hipStream_t st) {
  hipLaunchKernelGGL(k_count_tables, dim3(1), dim3(kBlock), 0, st);
  return hipGetLastError();
}

// Count waves with work that keep the device busy (four per SIMD).
constexpr int64_t kCountWaves = 4096;

void count4_tiling(const Plan& p, VolArgs* a, bool split) {
  const int ty = p.T == K_F32 ? count_ty<float>() : count_ty<double>();
  const int64_t cx = a->g.nx - 1, cy = a->g.ny - 1;
  a->czl = 0;
  if (cx <= 0 || cy <= 0 || a->g.nzseg <= 0) {
    a->cnbx = a->cnby = 0;
    a->cntiles = 0;
    return;
  }
  a->cnby = (int)((cy + ty - 1) / ty);
  // waves per x span: double them while the volume gives fewer than
  // kCountWaves waves and each still gets whole z blocks of its own
  const int64_t spans = (cx + kWave - 2) / (kWave - 1);
  const int64_t work = spans * a->cnby * a->g.nzseg;
  const int zblocks = (a->g.zseg + (1 << zmask_shift(a->g.zseg)) - 1) >> zmask_shift(a->g.zseg);
  int czl = 0;
  while (split && czl < 2 && (work << (czl + 1)) <= kCountWaves && (1 << (czl + 1)) <= zblocks) ++czl;
  a->czl = czl;
  const int per = (kWaves >> czl) * (kWave - 1);  // x cells per count tile
  a->cnbx = (int)((cx + per - 1) / per);
  a->cntiles = (int64_t)a->cnbx * a->cnby * a->g.nzseg;
}

template <typename T, int METHOD>
static hipError_t count4_t(const VolArgs& a, const VolArgs* vols, const int32_t* cmap,
                           int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  constexpr int TY = count_ty<T>();
  if (vols) {
    const auto k = k_count4_b<T, METHOD, TY, count_zu<T>()>;
    hipLaunchKernelGGL(k, dim3(persistent_grid((const void*)k, nblocks)), dim3(kBlock), 0, st,
                       vols, cmap, gbase, nblocks, ctr);
  } else
    hipLaunchKernelGGL((k_count4<T, METHOD, TY, count_zu<T>()>), dim3((unsigned)a.cntiles),
                       dim3(kBlock), 0, st, a);
  return hipGetLastError();
}

static hipError_t count4_any(const Plan& p, const VolArgs& a, const VolArgs* vols,
                             const int32_t* cmap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                             hipStream_t st) {
#define MSH_ARGS a, vols, cmap, gbase, nblocks, ctr, st
  if (p.T == K_F32) {
    if (p.method == MC) return count4_t<float, MC>(MSH_ARGS);
    if (p.method == MT) return count4_t<float, MT>(MSH_ARGS);
    return count4_t<float, SN>(MSH_ARGS);
  }
  if (p.method == MC) return count4_t<double, MC>(MSH_ARGS);
  if (p.method == MT) return count4_t<double, MT>(MSH_ARGS);
  return count4_t<double, SN>(MSH_ARGS);
#undef MSH_ARGS
}

hipError_t launch_count4(const Plan& p, const VolArgs& a, hipStream_t st) {
  if (a.cntiles == 0) return hipSuccess;
  return count4_any(p, a, nullptr, nullptr, 0, 0, nullptr, st);
}

hipError_t launch_count4_batch(const Plan& p, const VolArgs* vols, const int32_t* cmap,
                               int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  return count4_any(p, VolArgs{}, vols, cmap, gbase, nblocks, ctr, st);
}

// Tile shape of the Marching Cubes emit: z chunks of 16 cells and one cell
// row along y per block (on 1024^3 value noise this beat 32-cell chunks,
// whose larger LDS brick leaves fewer blocks per CU, and two y rows per
// block, likewise).
int mc4_variant(const Plan& p, VolArgs* a) {
  const int zc = 16;
  const int64_t esz = p.T == K_F32 ? 4 : 8;
  // the brick's DMA offsets are 32-bit and span ZC+1 planes
  if ((int64_t)(zc + 1) * a->g.nx * a->g.ny * esz >= (1ll << 31)) return -1;
  a->nti = (a->g.nx - 1 + kEmitTX - 1) / kEmitTX;
  a->ntj = a->g.ny - 1;
  a->ntiles = (int64_t)a->nti * a->ntj * a->g.nzseg;
  // 16-byte DMA needs 16-byte aligned sample rows
  const bool vec = ((uintptr_t)a->f % 16 == 0) && (((int64_t)a->g.nx * esz) % 16 == 0);
  const int cxclass = a->cxmask == 0 ? 1 : a->cxmask == 7 ? 2 : 0;
  return (vec ? 0 : 1) | (cxclass << 1);
}

template <typename T, typename M, typename Out, int ZC, int CY, int CXM>
static hipError_t emit_mc4_z(bool vec, const VolArgs& a, const VolArgs* vols,
                             const int32_t* emap, int64_t gbase, int64_t nblocks,
                             hipStream_t st) {
  if (vols) {
    if (vec)
      hipLaunchKernelGGL((k_emit_mc4_b<T, M, Out, ZC, CY, true, CXM>), dim3((unsigned)nblocks),
                         dim3(kBlock), 0, st, vols, emap, gbase);
    else
      hipLaunchKernelGGL((k_emit_mc4_b<T, M, Out, ZC, CY, false, CXM>), dim3((unsigned)nblocks),
                         dim3(kBlock), 0, st, vols, emap, gbase);
  } else {
    if (vec)
      hipLaunchKernelGGL((k_emit_mc4<T, M, Out, ZC, CY, true, CXM>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
    else
      hipLaunchKernelGGL((k_emit_mc4<T, M, Out, ZC, CY, false, CXM>), dim3((unsigned)a.ntiles),
                         dim3(kBlock), 0, st, a);
  }
  return hipGetLastError();
}

template <typename T, typename M, typename Out>
static hipError_t emit_mc4_t(int variant, const VolArgs& a, const VolArgs* vols,
                             const int32_t* emap, int64_t gbase, int64_t nblocks,
                             hipStream_t st) {
  const bool vec = (variant & 1) == 0;
  const int cxclass = (variant >> 1) & 3;
#define MSH_ARGS vec, a, vols, emap, gbase, nblocks, st
  if (cxclass == 1) return emit_mc4_z<T, M, Out, 16, 1, 0>(MSH_ARGS);
  if (cxclass == 2) return emit_mc4_z<T, M, Out, 16, 1, 7>(MSH_ARGS);
  return emit_mc4_z<T, M, Out, 16, 1, -1>(MSH_ARGS);
#undef MSH_ARGS
}

static hipError_t emit_mc4_any(const Plan& p, int variant, const VolArgs& a, const VolArgs* vols,
                               const int32_t* emap, int64_t gbase, int64_t nblocks,
                               hipStream_t st) {
  if (p.T == K_F32 && p.M == K_F32 && p.Out == K_F32)
    return emit_mc4_t<float, float, float>(variant, a, vols, emap, gbase, nblocks, st);
  if (p.T == K_F32 && p.M == K_F32)
    return emit_mc4_t<float, float, double>(variant, a, vols, emap, gbase, nblocks, st);
  if (p.T == K_F32) return emit_mc4_t<float, double, double>(variant, a, vols, emap, gbase, nblocks, st);
  return emit_mc4_t<double, double, double>(variant, a, vols, emap, gbase, nblocks, st);
}

hipError_t launch_emit_mc4(const Plan& p, int variant, const VolArgs& a, hipStream_t st) {
  if (a.ntiles == 0) return hipSuccess;
  return emit_mc4_any(p, variant, a, nullptr, nullptr, 0, 0, st);
}

hipError_t launch_emit_mc4_batch(const Plan& p, int variant, const VolArgs* vols,
                                 const int32_t* emap, int64_t gbase, int64_t nblocks,
                                 hipStream_t st) {
  if (nblocks == 0) return hipSuccess;
  return emit_mc4_any(p, variant, VolArgs{}, vols, emap, gbase, nblocks, st);
}

}  // namespace msh
