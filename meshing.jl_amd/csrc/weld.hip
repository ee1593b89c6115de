// weld.hip -- exact-coordinate vertex welding of a triangle mesh on the
// device: the opt-in step after Marching Cubes, whose cubes each emit their
// own copy of a vertex they share (src/algorithmtypes.jl:17-19).
//
// Two vertices weld when their three coordinates have identical bit
// patterns.  The welded vertices keep their order of first appearance and
// the faces are renumbered (1-based) to them, so the result is what a
// first-appearance `unique` over the vertex tuples gives on the host.
//
//   k_weld_bounds   (large tables only) one block sorts 4096 x coordinates
//                   sampled evenly over the vertices and keeps their
//                   quantiles as the bounds of the table's 64 x regions
//   k_weld_insert   open-addressing hash table over the coordinate bits,
//                   first within each block's 512 vertices in LDS, then
//                   over the whole mesh for the block's first appearances:
//                   the first vertex of a key claims a slot (CAS from -1),
//                   equal keys then lower it to their smallest index
//                   (atomic min, only when smaller than what the slot
//                   showed), so every slot ends holding the first appearance
//                   of its key; each vertex notes its slot
//   k_weld_rep      per vertex: its key's first appearance (one table load
//                   per local key); flag = first itself; per tile of 4096
//                   vertices: its count of flags
//   k_weld_tiles    one block: exclusive prefix of the tile counts, total =
//                   welded count
//   k_weld_pos      exclusive prefix of the flags = new index of each first
//                   appearance (reduce-then-scan: no block waits on another)
//   k_weld_compact  first appearances to their new slots
//   k_weld_faces    face indices renumbered in place to the new index of
//                   their vertex's first appearance (out-of-range indices
//                   are counted and left as they are)
//
// A key's home slot in a large table is its x region's block of slots plus
// the hash's low bits inside it.  Vertices of a mesh are emitted cell by
// cell with x outermost, so the blocks in flight at any moment work in one
// or two of the 64 regions rather than all over the table (on the 272 M
// vertex Marching Cubes mesh of 1024^3 value noise: Float64 7% faster,
// Float32 level; 512 regions and none were slower); the regions come from
// quantiles of the x values, so each takes a like share of the keys whatever
// their distribution, and keys whose x equals several bounds (a plane
// x = const) spread over those regions by the hash.  Placement only changes
// where keys sit: the table still ends holding each key's smallest index,
// and a probe that runs past its region goes on into the next one.
//
// Every probe loop ends: each table has more slots than keys (the global
// one at least 1.25 times the vertices, the local one twice its window),
// slots never return to empty, and a probe stops at an empty slot or at its
// own key.
#include "device.h"

namespace msh {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t h) {
  // SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable
  // pseudorandom number generators", OOPSLA 2014)
  h ^= h >> 30;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 27;
  h *= 0x94D049BB133111EBull;
  h ^= h >> 31;
  return h;
}

template <typename W>
__device__ __forceinline__ uint64_t key_hash(W x, W y, W z) {
  uint64_t h = mix64((uint64_t)x + 0x9E3779B97F4A7C15ull);
  h = mix64(h ^ (uint64_t)y);
  return mix64(h ^ (uint64_t)z);
}

__global__ void k_fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// An IEEE bit pattern as an unsigned integer whose order is the numeric
// order (-0 before +0, NaNs beyond the infinities of their sign).
__device__ __forceinline__ uint64_t ordered(uint32_t b) {
  return (b & 0x80000000u) ? (uint64_t)(uint32_t)~b : (uint64_t)(b | 0x80000000u);
}
__device__ __forceinline__ uint64_t ordered(uint64_t b) {
  return (b >> 63) ? ~b : b | (1ull << 63);
}

constexpr int kWeldSample = 4096;     // x coordinates sorted for the region bounds
constexpr int kWeldRegions = 64;      // regions of a large table (measured: 64 beat 512 and none)
constexpr int kWeldSortBlock = 1024;

// One block: bounds[r - 1] = the (r / nreg) quantile of the sampled x
// coordinates (as ordered integers), r = 1 .. nreg - 1 (bitonic sort in LDS).
template <typename W>
__global__ __launch_bounds__(kWeldSortBlock) void k_weld_bounds(const W* __restrict__ v,
                                                               int64_t nv, int nreg,
                                                               uint64_t* __restrict__ bounds) {
  __shared__ uint64_t s[kWeldSample];
  for (int k = threadIdx.x; k < kWeldSample; k += kWeldSortBlock)
    s[k] = ordered(v[3 * ((int64_t)k * nv / kWeldSample)]);
  __syncthreads();
  for (int size = 2; size <= kWeldSample; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < kWeldSample; k += kWeldSortBlock) {
        const int p = k ^ stride;
        if (p > k) {
          const uint64_t a = s[k], b = s[p];
          if ((a > b) == ((k & size) == 0)) {  // ascending runs where k & size is 0
            s[k] = b;
            s[p] = a;
          }
        }
      }
      __syncthreads();
    }
  for (int r = threadIdx.x + 1; r < nreg; r += kWeldSortBlock)
    bounds[r - 1] = s[(int64_t)r * kWeldSample / nreg];
}

// Home slot of a key with hash h: region (from x and the nreg - 1 sorted
// bounds in bnd) in the bits from rbits up, the hash's low bits below.
template <typename W>
__device__ __forceinline__ uint64_t home_slot(W x, uint64_t h, const uint64_t* bnd, int nreg,
                                              int rbits) {
  const uint64_t low = h & ((1ull << rbits) - 1);
  if (nreg == 1) return low;
  const uint64_t o = ordered(x);
  int a = 0, b = nreg - 1;  // c = bounds below o
  while (a < b) {
    const int m = (a + b) >> 1;
    if (bnd[m] < o) a = m + 1; else b = m;
  }
  const int c = a;
  uint64_t r = c;
  if (c < nreg - 1 && bnd[c] == o) {  // o is a bound: e = bounds at or below o
    b = nreg - 1;
    while (a < b) {
      const int m = (a + b) >> 1;
      if (bnd[m] <= o) a = m + 1; else b = m;
    }
    r += (h >> 32) % (uint64_t)(a - c + 1);
  }
  return (r << rbits) | low;
}

// Global insert of vertex i with key (x, y, z), probing from slot h:
// returns its slot.
template <typename W>
__device__ __forceinline__ uint32_t global_insert(const W* __restrict__ v, int64_t i, W x, W y,
                                                  W z, uint64_t h, unsigned long long* table,
                                                  uint64_t mask) {
  while (true) {
    // a plain load first: a CAS on a taken slot measured slower
    unsigned long long cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
    if (cur == ~0ull) {
      cur = atomicCAS(&table[h], ~0ull, (unsigned long long)i);
      if (cur == ~0ull) break;  // claimed
    }
    // every index a slot ever holds has the slot's key, so this compare
    // does not depend on which of them it reads
    if (v[3 * cur] == x && v[3 * cur + 1] == y && v[3 * cur + 2] == z) {
      // a slot only ever decreases: above what it showed, i cannot lower it
      if ((unsigned long long)i < cur) atomicMin(&table[h], (unsigned long long)i);
      break;
    }
    h = (h + 1) & mask;
  }
  return (uint32_t)h;
}

constexpr int kWeldLT = 512;           // vertices per block: the local pass's window
constexpr int kWeldLS = 2 * kWeldLT;   // local hash slots

// Copies of a vertex sit close together in extraction order (the cubes
// around an edge are neighbours in the same or the next cell row), so each
// block first welds its 512 consecutive vertices in LDS and sends only
// the first appearance of each local key to the global table; each vertex
// notes that local first appearance, whose slot k_weld_rep reads once for
// all of them.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_weld_insert(const W* __restrict__ v, int64_t nv,
                                                        unsigned long long* table,
                                                        uint64_t mask,
                                                        const uint64_t* __restrict__ bounds,
                                                        int nreg, int rbits,
                                                        uint32_t* __restrict__ slot,
                                                        uint16_t* __restrict__ lrep) {
  __shared__ W key[3 * kWeldLT];
  __shared__ uint32_t lt[kWeldLS];  // local slot -> smallest local index of its key
  __shared__ uint64_t bnd[kWeldRegions - 1];
  constexpr uint32_t kEmpty = 0xFFFFFFFFu;
  constexpr int PER = kWeldLT / kBlock;
  const int64_t base = (int64_t)blockIdx.x * kWeldLT;
  const int n = nv - base < kWeldLT ? (int)(nv - base) : kWeldLT;
  for (int k = threadIdx.x; k < kWeldLS; k += kBlock) lt[k] = kEmpty;
  for (int k = threadIdx.x; k < nreg - 1; k += kBlock) bnd[k] = bounds[k];
  for (int k = threadIdx.x; k < 3 * n; k += kBlock) key[k] = v[3 * base + k];
  __syncthreads();
  uint32_t ls[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = threadIdx.x + q * kBlock;
    ls[q] = 0;
    if (j >= n) continue;
    const W x = key[3 * j], y = key[3 * j + 1], z = key[3 * j + 2];
    uint32_t h = (uint32_t)key_hash(x, y, z) & (kWeldLS - 1);
    while (true) {
      uint32_t cur = __hip_atomic_load(&lt[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (cur == kEmpty) {
        cur = atomicCAS(&lt[h], kEmpty, (uint32_t)j);
        if (cur == kEmpty) break;
      }
      if (key[3 * cur] == x && key[3 * cur + 1] == y && key[3 * cur + 2] == z) {
        if ((uint32_t)j < cur) atomicMin(&lt[h], (uint32_t)j);
        break;
      }
      h = (h + 1) & (kWeldLS - 1);
    }
    ls[q] = h;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = threadIdx.x + q * kBlock;
    if (j >= n) continue;
    const uint32_t l = lt[ls[q]];
    lrep[base + j] = (uint16_t)l;
    if (l == (uint32_t)j) {
      const W x = key[3 * j], y = key[3 * j + 1], z = key[3 * j + 2];
      const uint64_t h = home_slot(x, key_hash(x, y, z), bnd, nreg, rbits);
      slot[base + j] = global_insert(v, base + j, x, y, z, h, table, mask);
    }
  }
}

constexpr int kWeldTile = 4096;                 // vertices per prefix tile
constexpr int kWeldPer = kWeldTile / kBlock;     // per thread in k_weld_pos
static_assert(kWeldLT % kBlock == 0 && kWeldTile % kWeldLT == 0 && kWeldLT <= 65536,
              "windows of whole waves inside one prefix tile, 16-bit local indices");

// Same windows as k_weld_insert: the local first appearances read their
// key's smallest index from the table, the others take it from them.
__global__ __launch_bounds__(kBlock) void k_weld_rep(int64_t nv,
                                                     const int64_t* __restrict__ table,
                                                     const uint32_t* __restrict__ slot,
                                                     const uint16_t* __restrict__ lrep,
                                                     int64_t* __restrict__ rep,
                                                     uint32_t* __restrict__ flag,
                                                     uint32_t* __restrict__ tilecnt) {
  __shared__ int64_t res[kWeldLT];
  constexpr int PER = kWeldLT / kBlock;
  const int64_t base = (int64_t)blockIdx.x * kWeldLT;
  const int n = nv - base < kWeldLT ? (int)(nv - base) : kWeldLT;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = threadIdx.x + q * kBlock;
    if (j < n && lrep[base + j] == j) res[j] = table[slot[base + j]];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int j = threadIdx.x + q * kBlock;
    const int64_t i = base + j;
    bool first = false;
    if (j < n) {
      const int64_t r = res[lrep[i]];  // the smallest index of i's key
      rep[i] = r;
      first = r == i;
      flag[i] = first ? 1u : 0u;
    }
    // a wave's 64 vertices here are consecutive and lie in one tile
    const uint64_t b = __ballot(first);
    if (lane_id() == 0 && b != 0) atomicAdd(&tilecnt[i / kWeldTile], (uint32_t)__popcll(b));
  }
}

// Exclusive prefix of the flags in three steps: the tile counts (above),
// their prefix in one block (here), then each tile's own (k_weld_pos).
__global__ __launch_bounds__(1024) void k_weld_tiles(const uint32_t* __restrict__ tilecnt,
                                                     int64_t ntiles, int64_t* __restrict__ tileoff,
                                                     int64_t* __restrict__ total) {
  __shared__ int64_t s_w[1024 / kWave];
  const int t = threadIdx.x, lane = lane_id(), wv = t / kWave;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t lo = t * per, hi = lo + per < ntiles ? lo + per : ntiles;
  int64_t sum = 0;
  for (int64_t k = lo; k < hi; ++k) sum += tilecnt[k];
  const int64_t incl = (int64_t)wave_incl_scan64((uint64_t)sum);
  if (lane == kWave - 1) s_w[wv] = incl;
  __syncthreads();
  int64_t before = 0;
  for (int q = 0; q < wv; ++q) before += s_w[q];
  int64_t run = before + incl - sum;
  for (int64_t k = lo; k < hi; ++k) {
    tileoff[k] = run;
    run += tilecnt[k];
  }
  if (t == 1023) *total = before + incl;
}

// One tile per block in kWeldPer rounds of 256 consecutive vertices, so
// every load and store of a round is one coalesced stretch.
__global__ __launch_bounds__(kBlock) void k_weld_pos(const uint32_t* __restrict__ flag,
                                                     int64_t nv,
                                                     const int64_t* __restrict__ tileoff,
                                                     int64_t* __restrict__ pos) {
  __shared__ uint32_t s_w[kWeldPer][kWaves];
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  const int64_t base = (int64_t)blockIdx.x * kWeldTile + threadIdx.x;
  uint32_t f[kWeldPer], incl[kWeldPer];
#pragma unroll
  for (int q = 0; q < kWeldPer; ++q) {
    const int64_t i = base + q * kBlock;
    f[q] = i < nv ? flag[i] : 0u;
    incl[q] = wave_scan_dpp(f[q]);
    if (lane == kWave - 1) s_w[q][wv] = incl[q];
  }
  __syncthreads();
  int64_t run = tileoff[blockIdx.x];
#pragma unroll
  for (int q = 0; q < kWeldPer; ++q) {
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      before += w < wv ? s_w[q][w] : 0u;
      all += s_w[q][w];
    }
    const int64_t i = base + q * kBlock;
    if (i < nv) pos[i] = run + before + incl[q] - f[q];
    run += all;
  }
}

template <typename W>
__global__ void k_weld_compact(const W* __restrict__ v, int64_t nv,
                               const int64_t* __restrict__ pos,
                               const uint32_t* __restrict__ flag, W* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv || !flag[i]) return;
  const int64_t o = pos[i];  // a first appearance's new index
  out[3 * o] = v[3 * i];
  out[3 * o + 1] = v[3 * i + 1];
  out[3 * o + 2] = v[3 * i + 2];
}

__global__ void k_weld_faces(int64_t* __restrict__ f, int64_t n,
                             const int64_t* __restrict__ rep, const int64_t* __restrict__ pos,
                             int64_t nv, unsigned long long* __restrict__ bad) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t j = f[k];
  if (j >= 1 && j <= nv)
    f[k] = pos[rep[j - 1]] + 1;  // the new index of the vertex's first appearance
  else
    atomicAdd(bad, 1ull);
}

unsigned grid_of(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename W>
hipError_t weld_t(const WeldArgs& a, hipStream_t st) {
  const W* v = (const W*)a.verts;
  hipLaunchKernelGGL(k_fill_i64, dim3(1), dim3(kBlock), 0, st, a.bad, (int64_t)1, (int64_t)0);
  if (a.nverts > 0) {
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_of(a.nslots)), dim3(kBlock), 0, st, a.table,
                       a.nslots, (int64_t)-1);
    // slots (4 bytes per vertex) and local first appearances (2 bytes) live
    // in pos until k_weld_pos overwrites it
    uint32_t* slot = (uint32_t*)a.pos;
    uint16_t* lrep = (uint16_t*)(slot + a.nverts);
    const unsigned nwin = (unsigned)((a.nverts + kWeldLT - 1) / kWeldLT);
    // x regions only where the table outgrows the caches (2^23 slots,
    // 64 MiB, and up)
    int lg = 0;
    while (((int64_t)1 << lg) < a.nslots) ++lg;
    const int nreg = lg >= 23 ? kWeldRegions : 1;
    int rbits = lg;
    for (int r = nreg; r > 1; r >>= 1) --rbits;
    if (nreg > 1)
      hipLaunchKernelGGL((k_weld_bounds<W>), dim3(1), dim3(kWeldSortBlock), 0, st, v, a.nverts,
                         nreg, a.bounds);
    hipLaunchKernelGGL((k_weld_insert<W>), dim3(nwin), dim3(kBlock), 0, st, v, a.nverts,
                       (unsigned long long*)a.table, (uint64_t)(a.nslots - 1),
                       (const uint64_t*)a.bounds, nreg, rbits, slot, lrep);
    const int64_t ntiles = (a.nverts + kWeldTile - 1) / kWeldTile;
    hipLaunchKernelGGL(k_fill_i64, dim3(grid_of((ntiles + 1) / 2)), dim3(kBlock), 0, st,
                       (int64_t*)a.tilecnt, (ntiles + 1) / 2, (int64_t)0);
    hipLaunchKernelGGL(k_weld_rep, dim3(nwin), dim3(kBlock), 0, st, a.nverts,
                       (const int64_t*)a.table, (const uint32_t*)slot, (const uint16_t*)lrep,
                       a.rep, a.flag, a.tilecnt);
    hipLaunchKernelGGL(k_weld_tiles, dim3(1), dim3(1024), 0, st, (const uint32_t*)a.tilecnt,
                       ntiles, a.tileoff, a.total);
    hipLaunchKernelGGL(k_weld_pos, dim3((unsigned)ntiles), dim3(kBlock), 0, st,
                       (const uint32_t*)a.flag, a.nverts, (const int64_t*)a.tileoff, a.pos);
    hipLaunchKernelGGL((k_weld_compact<W>), dim3(grid_of(a.nverts)), dim3(kBlock), 0, st, v,
                       a.nverts, (const int64_t*)a.pos, (const uint32_t*)a.flag,
                       (W*)a.verts_out);
  }
  if (a.nfaces > 0)
    hipLaunchKernelGGL(k_weld_faces, dim3(grid_of(3 * a.nfaces)), dim3(kBlock), 0, st, a.faces,
                       3 * a.nfaces, (const int64_t*)a.rep, (const int64_t*)a.pos, a.nverts,
                       (unsigned long long*)a.bad);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_weld(const WeldArgs& a, hipStream_t st) {
  return a.out_kind == K_F32 ? weld_t<uint32_t>(a, st) : weld_t<uint64_t>(a, st);
}

}  // namespace msh
