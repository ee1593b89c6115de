// device.h -- device helpers shared by the kernels of libmeshing_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include "internal.h"

namespace msh {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

// Corner offsets.  Marching Cubes order (src/marching_cubes.jl:42-49):
// (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1).
__device__ __forceinline__ int mc_dx(int c) { return (0x66 >> c) & 1; }
__device__ __forceinline__ int mc_dy(int c) { return (0xCC >> c) & 1; }
__device__ __forceinline__ int mc_dz(int c) { return (0xF0 >> c) & 1; }
// Marching Tetrahedra order (src/marching_tetrahedra.jl:146-153):
// (0,0,0) (0,1,0) (1,1,0) (1,0,0) (0,0,1) (0,1,1) (1,1,1) (1,0,1).
__device__ __forceinline__ int mt_dx(int c) { return (0xCC >> c) & 1; }
__device__ __forceinline__ int mt_dy(int c) { return (0x66 >> c) & 1; }
__device__ __forceinline__ int mt_dz(int c) { return (0xF0 >> c) & 1; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
// Wave index inside the block, in a scalar register (it is the same for all
// lanes; saying so keeps address math that depends on it scalar).
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x / kWave));
}

// Make this wave's LDS writes visible to its own later LDS reads.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint64_t t = __shfl_up(v, d, kWave);
    if (lane >= d) v += t;
  }
  return v;
}

// Workgroup barrier ordering LDS only: it does not wait for this wave's
// outstanding global stores, which __syncthreads() would.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The OR over the block of every thread's 64-bit value and of its flag,
// with one barrier (`slot`: 2 * kWaves words of LDS, free until every wave
// has passed the barrier that starts the next tile): a tile's z-block
// activity mask from its rows' masks, and whether any row writes.
__device__ __forceinline__ uint64_t tile_zmask(uint64_t m, bool flag, uint64_t* slot,
                                               bool* any) {
  for (int d = kWave / 2; d >= 1; d >>= 1) m |= __shfl_xor(m, d, kWave);
  const bool f = __ballot(flag) != 0;
  if ((threadIdx.x & (kWave - 1)) == 0) {
    slot[threadIdx.x / kWave] = m;
    slot[kWaves + threadIdx.x / kWave] = f ? 1u : 0u;
  }
  lds_barrier();
  uint64_t t = 0, fl = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    t |= slot[w];
    fl |= slot[kWaves + w];
  }
  // the same in every lane: keep it (and the chunk walk it steers) scalar
  *any = __builtin_amdgcn_readfirstlane((int)fl) != 0;
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(t >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
}

// The first chunk after the one at zc (chunks of zcl cells from zs, which
// never straddle a z block of 1 << zsh cells) whose block is marked in
// tmask, or ze when there is none.
__device__ __forceinline__ int next_zchunk(int zc, int zcl, int zs, int ze, int zsh,
                                           uint64_t tmask) {
  const int nz = zc + zcl;
  if (nz >= ze) return ze;
  const int b = (nz - zs) >> zsh;
  if ((tmask >> b) & 1) return nz;
  const uint64_t rest = b < 63 ? tmask >> (b + 1) : 0ull;
  if (rest == 0) return ze;
  const int zn = zs + ((b + 1 + __builtin_ctzll(rest)) << zsh);
  return zn < ze ? zn : ze;
}

// Lane i gets lane i+1's value (DPP wave_shl:1); lane 63 gets `last`.
__device__ __forceinline__ uint32_t next_lane(uint32_t v, uint32_t last) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)v, 0x130, 0xF, 0xF, false);
}

// Blocks of one launch are dealt round-robin to the 8 XCDs (each with its
// own L2).  Renumber so that each XCD gets a contiguous range of tiles:
// neighbouring tiles (which share halo samples) then meet in the same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nb) {
  const int64_t xcd = b & 7, local = b >> 3;
  const int64_t per = nb >> 3, rem = nb & 7;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}



// Tile sources of the kernels' tile loops.  A single-volume launch runs one
// tile per block (OneTile).  A batched launch is persistent: as many blocks
// as fit on the device at once, each taking tiles from a work queue until
// the batch's tiles are gone (TileQueue), so the per-block set-up (lookup
// tables in LDS) is paid once per block rather than once per tile, and
// volumes of different sizes balance across the CUs.  The launch's tiles are
// cut into 8 contiguous ranges, one per XCD (block b starts on XCD b % 8);
// a block takes tiles from its own XCD's range first, so tiles that are
// neighbours in the batch (and share cache lines at their edges) meet in
// that XCD's L2, and then helps the other ranges in turn.  Each range has
// its own counter (ctr[0..7], zero at launch).
struct OneTile {
  static constexpr bool kOne = true;  // the kernels call the tile directly
  const VolArgs* p;
  int64_t t;
  bool done;
  __device__ bool operator()(VolArgs& a, int64_t& tile) {
    if (done) return false;
    done = true;
    a = *p;
    tile = t;
    return true;
  }
};

template <bool EMIT>
struct TileQueue {
  static constexpr bool kOne = false;
  const VolArgs* __restrict__ vols;
  const int32_t* __restrict__ map;  // block of the batch -> volume
  int64_t gbase, n;                 // the launch's first block of the batch, its tiles
  uint32_t* __restrict__ ctr;       // 8 counters
  int k;                            // ranges exhausted so far
  __device__ bool operator()(VolArgs& a, int64_t& tile) {
    __shared__ int64_t slot;
    const int q0 = (int)(blockIdx.x & 7);
    while (k < 8) {
      const int q = (q0 + k) & 7;
      const int64_t per = n >> 3, rem = n & 7;
      const int64_t lo = q * per + (q < rem ? q : rem), cnt = per + (q < rem ? 1 : 0);
      lds_barrier();  // every wave has read the previous slot
      if (threadIdx.x == 0) slot = (int64_t)atomicAdd(&ctr[q], 1u);
      lds_barrier();
      // uniform across the block: keep it (and the descriptor loads) scalar
      const uint64_t raw = (uint64_t)slot;
      const int64_t i = (int64_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(
                                      (int)(raw >> 32)) << 32) |
                                  (uint32_t)__builtin_amdgcn_readfirstlane((int)raw));
      if (i < cnt) {
        const int64_t b = gbase + lo + i;
        a = vols[map[b]];
        tile = b - (EMIT ? a.eblk0 : a.cblk0);
        return true;
      }
      ++k;
    }
    return false;
  }
};

// Inclusive scan over the 64 lanes with DPP: Kogge-Stone within each row of
// 16 lanes, then row broadcasts of lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

// One axis of p1 .+ mu .* (p2 .- p1) (src/marching_cubes.jl:100-104) with
// p1, p2 of the axis element type; the result type is promote(M, CX).
template <typename M>
__device__ __forceinline__ double lerp_axis(bool cx32, M mu, double p1, double p2) {
  if (cx32) {
    const float a = (float)p1, b = (float)p2;
    return (double)(a + mu * (b - a));
  }
  return (double)(p1 + mu * (p2 - p1));
}

// One axis of vertPos: base .+ (c1*b .+ c2*a) .* (next - base)
// (src/marching_tetrahedra.jl:49-54), with w = c1*b + c2*a of type A.
template <typename A>
__device__ __forceinline__ double mt_axis(bool cx32, A w, double lo, double hi) {
  if (cx32) {
    const float l = (float)lo, h = (float)hi;
    const float d = h - l;
    return (double)(l + w * d);
  }
  const double d = hi - lo;
  return (double)(lo + w * d);
}

// Julia's min/max of floats: NaN-propagating, -0.0 < +0.0.
template <typename C>
__device__ __forceinline__ C jl_min(C x, C y) {
  if (isnan(x)) return x;
  if (isnan(y)) return y;
  return signbit(x - y) ? x : y;
}
template <typename C>
__device__ __forceinline__ C jl_max(C x, C y) {
  if (isnan(x)) return x;
  if (isnan(y)) return y;
  return signbit(x - y) ? y : x;
}


// 4-bit code of the sample plane under a cell's lower face, in the corner
// order of METHOD.  p points at sample (ci, cj) of that plane.
template <int METHOD, typename T, int SX>
__device__ __forceinline__ uint32_t face_code(const T* p, T thr) {
  const uint32_t b00 = p[0] < thr, b10 = p[1] < thr, b11 = p[SX + 1] < thr, b01 = p[SX] < thr;
  if (METHOD == MT) return b00 | (b01 << 1) | (b11 << 2) | (b10 << 3);
  return b00 | (b10 << 1) | (b11 << 2) | (b01 << 3);
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}


// Inclusive scan over independent segments of SEG lanes (16, 32 or 64) with
// DPP: Kogge-Stone within each row of 16 lanes, then row broadcasts.
template <int SEG>
__device__ __forceinline__ uint32_t seg_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
  if (SEG >= 32)
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
  if (SEG >= 64)
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
  return v;
}

// Raw buffer resource over `bytes` bytes at p (32-bit data format, so
// loads past `bytes` return 0 instead of touching memory).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, 0x00020000);
}

// Natural nibble order -> Marching Cubes corner order (src/marching_cubes.jl:42-49):
// swap bits 2 and 3 of each nibble.
__device__ __forceinline__ uint32_t nat_to_mc(uint32_t c) {
  return (c & 0x33u) | ((c & 0x44u) << 1) | ((c & 0x88u) >> 1);
}
// Natural nibble order -> Marching Tetrahedra corner order
// (src/marching_tetrahedra.jl:146-153): (x,y) (x,y+1) (x+1,y+1) (x+1,y).
__device__ __forceinline__ uint32_t nat_to_mt(uint32_t c) {
  return (c & 0x11u) | ((c & 0x44u) >> 1) | ((c & 0x88u) >> 1) | ((c & 0x22u) << 2);
}

// One LDS DMA instruction: each lane copies BYTES bytes from the buffer at
// byte offset `off` to lds + lane*BYTES.  (The 16-byte form exists only in
// the gfx950 device compile; the host compile never runs kernels.)
template <int BYTES>
__device__ __forceinline__ void dma_to_lds(__amdgpu_buffer_rsrc_t r, uint32_t* lds, uint32_t off) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (BYTES == 16)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (void*)lds, 16, off, 0, 0, 0);
  else
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (void*)lds, 4, off, 0, 0, 0);
#endif
}

// lean.hip: the count kernel of every method and the Marching Cubes emit kernel.
void count4_tiling(const Plan& p, VolArgs* a, bool split);
hipError_t launch_count4(const Plan& p, const VolArgs& a, hipStream_t st);
hipError_t launch_count4_batch(const Plan& p, const VolArgs* vols, const int32_t* cmap,
                               int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st);
// -1: volume too large for k_emit_mc4; else bit 0: dword DMA (not 16-byte),
// bits 1-2: axis types (1 all Float64, 2 all Float32, 0 mixed), bits 3-4:
// tile shape.  Fills the emit tiling.
int mc4_variant(const Plan& p, VolArgs* a);
hipError_t launch_emit_mc4(const Plan& p, int variant, const VolArgs& a, hipStream_t st);
// mc5.hip: the Marching Cubes emit kernel with lanes along x.  -1: volume
// too large; else bit 0: dword DMA, bits 1-2: axis types.  Fills the tiling.
int mc5_variant(const Plan& p, VolArgs* a);
hipError_t launch_emit_mc5(const Plan& p, int variant, const VolArgs& a, hipStream_t st);
hipError_t launch_emit_mc5_batch(const Plan& p, int variant, const VolArgs* vols,
                                 const int32_t* emap, int64_t gbase, int64_t nblocks,
                                 uint32_t* ctr, hipStream_t st);

// halo5.hip: the Marching Tetrahedra and Naive Surface Nets emit kernel
// with lanes along x.  -1: volume too large; else bit 0: dword DMA, bits
// 1-2: axis types.  Fills the emit tiling.
int h5_variant(const Plan& p, VolArgs* a);
hipError_t launch_emit_h5(const Plan& p, int variant, const VolArgs& a, hipStream_t st);
hipError_t launch_emit_h5_batch(const Plan& p, int variant, const VolArgs* vols,
                                const int32_t* emap, int64_t gbase, int64_t nblocks,
                                uint32_t* ctr, hipStream_t st);

// Blocks of `kernel` (256 threads, static LDS only) that fit on the current
// device at once, capped at `ntiles`: the grid of a persistent launch.
unsigned persistent_grid(const void* kernel, int64_t ntiles);

// tet.hip: the Marching Tetrahedra and Naive Surface Nets emit kernel.  -1:
// volume too large; else bit 0: dword DMA, bits 1-2: axis types.  Fills the
// emit tiling.
int halo_variant(const Plan& p, VolArgs* a);
hipError_t launch_emit_halo(const Plan& p, int variant, const VolArgs& a, hipStream_t st);
hipError_t launch_emit_halo_batch(const Plan& p, int variant, const VolArgs* vols,
                                  const int32_t* emap, int64_t gbase, int64_t nblocks,
                                  hipStream_t st);
hipError_t launch_emit_mc4_batch(const Plan& p, int variant, const VolArgs* vols,
                                 const int32_t* emap, int64_t gbase, int64_t nblocks,
                                 hipStream_t st);

}  // namespace msh
