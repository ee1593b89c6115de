// internal.h -- shared between the C ABI (capi.cpp) and the kernels
// (kernels.hip).  Not installed; the public interface is include/meshing_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/meshing_hip.h"

namespace msh {

enum Kind : int { K_INT64 = 0, K_F32 = 1, K_F64 = 2 };
enum Method : int { MC = 0, MT = 1, SN = 2 };

// Cells are (nx-1)*(ny-1)*(nz-1).  The emission order of the reference is
// x outermost, z innermost (src/marching_cubes.jl:40).  A "unit" is one
// z-row of cells (fixed x, y) cut into nzseg segments of zseg cells; units
// are numbered in emission order: u = (i*(ny-1) + j)*nzseg + s.
struct Geom {
  int32_t nx, ny, nz;
  int32_t zseg, nzseg;
  int64_t units;
};

// Arithmetic types of one extraction, derived from the Julia types of the
// arguments exactly as Julia's promote_type does (src/marching_cubes.jl:31,
// :100-104; src/marching_tetrahedra.jl:43-54,131).
struct Plan {
  int method;
  int T;       // eltype(sdf)
  int M;       // promote(typeof(iso), T): type of the interpolation weight
  int A;       // promote(M, typeof(eps)): type of the clamped MT weight
  int Out;     // float(FT)
  int cx[3];   // LinRange element type per axis
  double thr;  // smallest T value >= iso, so `v < iso` == `v < thr` exactly
  double isoM; // iso converted to M
  double epsA; // eps converted to A
  double omeA; // one(T) - eps, computed in promote(T, typeof(eps)), as A
  double a[3], b[3];  // axis endpoints in the axis element type
  // emit kernel family: 3 as 2 with half-depth bricks (Marching Cubes with
  // one element type throughout and Float64 axes; otherwise as 2), 2 lanes
  // along x (k_emit_mc5 / k_emit_h5), 1 lanes along z (k_emit_mc4 /
  // k_emit_halo), 0 the generic k_emit; -1 the default order (2, 1, 0,
  // falling back when a volume does not fit)
  int family = -1;
};

struct Workspace {
  uint32_t* cnt;      // per unit: vertices | faces << 16
  uint64_t* zmask;    // per count row group (zmask_at): bit k set when z
                      // block k of the segment holds a cell with outputs
                      // (zmask_shift); room for one per unit
  int64_t* voff;      // per unit: exclusive prefix of vertices
  int64_t* foff;      // per unit: exclusive prefix of faces
  uint64_t* tilesum;  // per scan tile: vertices | faces << 32
  int64_t* tileoff;   // per scan tile: exclusive prefixes (v, f)
  int64_t* totals;    // (0, nverts, 0, nfaces): the sizes of a single volume
  double* coords;     // xp[nx], yp[ny], zp[nz] as double
};

// cell rows along y per count tile (one lane marches TY rows)
template <typename T>
__host__ __device__ constexpr int count_ty() { return sizeof(T) == 4 ? 8 : 6; }

// Everything the kernels need to know about one volume.  Single-volume
// launches take it by value; batched launches read one entry per volume
// from an array in device memory.
struct VolArgs {
  const void* f;           // field, x fastest
  Geom g;
  double thr, iso;         // classification threshold in T, iso in M
  double eps, ome;         // MT: eps and one(T)-eps, in A
  int32_t cxmask;          // bit k: axis k coordinates are Float32
  int32_t method;
  double ca[3], cb[3];     // LinRange ends per axis
  double* coords;          // xp | yp | zp of this volume
  uint32_t* cnt;           // unit counts, indexed from ubase
  uint64_t* zmask;         // z-block activity masks per count row group, from ubase
  const int64_t* voff;     // exclusive unit prefixes, indexed from ubase
  const int64_t* foff;
  const int64_t* sz;       // (first vertex, nverts, first face, nfaces) of this
                           // volume in voff/foff numbering; writes past the
                           // counts are dropped
  void* verts;             // this volume's outputs
  int64_t* faces;
  int64_t ubase;           // first unit of this volume
  int32_t cnbx, cnby;      // count tiles along x and y
  int32_t czl;             // log2 of the waves that share a count tile's z segment
  int64_t cntiles;         // count tiles (x * y * z segments)
  int64_t cblk0;           // first count block of this volume in a batched launch
  int32_t nti, ntj;        // emit tiles along x and y
  int64_t ntiles;          // emit tiles
  int64_t eblk0;           // first emit block of this volume in a batched launch
};

// The z-block mask of cell row (gi, gj) in z segment s: the count keeps one
// per lane and count tile, i.e. per count_ty<T>() consecutive rows along y,
// so a row's mask may mark blocks of its neighbours too.
template <typename T>
__host__ __device__ inline uint64_t zmask_at(const VolArgs& a, int gi, int gj, int s) {
  return a.zmask[a.ubase + ((int64_t)gi * a.cnby + gj / count_ty<T>()) * a.g.nzseg + s];
}

// Size classes of the kernel-family choice (msh_autotune keeps one choice
// per method, element type and class): volumes of fewer than 2^20 cells
// (about 100^3), fewer than 2^23 (about 200^3), and larger.
constexpr int kSizeClasses = 3;
inline int size_class(const msh_field* f) {
  const double cells = (double)(f->nx > 1 ? f->nx - 1 : 0) * (double)(f->ny > 1 ? f->ny - 1 : 0) *
                       (double)(f->nz > 1 ? f->nz - 1 : 0);
  return cells < 1048576.0 ? 0 : cells < 8388608.0 ? 1 : 2;
}

// z blocks of the activity masks: 16 cells for segments up to 1024 cells,
// 32 up to 2048, 64 up to 4096 (so a mask has at most 64 blocks, and every
// emit chunk of 4, 8 or 16 cells lies inside one block); returns log2 of the
// block length.
__host__ __device__ inline int zmask_shift(int zseg) {
  return zseg <= 1024 ? 4 : zseg <= 2048 ? 5 : 6;
}

constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;

inline int64_t scan_tiles(int64_t units) { return (units + kScanTile - 1) / kScanTile; }
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Julia's promotion for one call (capi.cpp); on error returns an msh_status
// and, when err is not null, a message in *err.
int make_plan(std::string* err, const msh_field* f, const msh_params* p, Plan* out);
// Units of a volume: z rows cut into segments (whole: split only to keep
// units within 4096 cells).
Geom make_geom(const msh_field* f, int method, bool whole = false);
// Workspace of `units` units and `ncoords` coordinates placed at `base`
// (null: only the size); returns the bytes it takes.
size_t workspace_layout(int64_t units, int64_t ncoords, void* base, Workspace* w);

// VolArgs of a single extraction (sz = w.totals, ubase = 0), with the
// count and emit tilings filled in.
VolArgs make_vol(const Plan& p, const Geom& g, const void* field, const Workspace& w);
// Emit kernel variant of a volume (volumes of one batched emit launch share
// it) and its tiling; fills nti, ntj, ntiles.
int emit_variant(const Plan& p, VolArgs* a);
// the family (Plan::family) a variant belongs to
int variant_family(int variant);
// count tiles of a volume; `split`: waves may share a tile's z segment when
// the volume alone is too small to fill the device (not in a batch, whose
// persistent launch is kept full by the other volumes)
void count_tiling(const Plan& p, VolArgs* a, bool split);

hipError_t launch_coords(const Plan& p, const Geom& g, double* coords, hipStream_t st);
hipError_t launch_count(const Plan& p, const VolArgs& a, hipStream_t st);
// The LinRange coordinates of a volume's three axes, end to end, into out.
struct CoordArgs {
  double* out;
  int nx, ny, nz;
  double a[3], b[3];
  int mask;  // bit k: axis k in Float32
};
CoordArgs coord_args(const Plan& p, const Geom& g, double* coords);
// host: pinned host words that also receive the vertex and face totals
// (at [1] and [3]), or null; coords: coordinate tables the scan also fills
// (in its one-block tile scan, or block 0 of the fused apply) instead of a
// launch of their own, or null
hipError_t launch_scan(const Geom& g, const Workspace& w, hipStream_t st, int64_t* host = nullptr,
                       const CoordArgs* coords = nullptr);
hipError_t launch_emit(const Plan& p, const VolArgs& a, hipStream_t st);

// Batched launches over n volumes whose VolArgs are in device memory.
// Volumes of one count/emit launch share the kernel instantiation (same
// Plan types and emit variant); gbase is the launch's first block in the
// block->volume maps.
hipError_t launch_coords_batch(const VolArgs* vols, int n, hipStream_t st);
hipError_t launch_fill_maps(const VolArgs* vols, int n, int32_t* cmap, int32_t* emap,
                            hipStream_t st);
// Persistent count and emit launches (families 2 and 3) take their tiles
// from 8 work-queue counters at ctr, which must be zero at launch.
hipError_t launch_count_batch(const Plan& p, const VolArgs* vols, const int32_t* cmap,
                              int64_t gbase, int64_t nblocks, uint32_t* ctr, hipStream_t st);
hipError_t launch_sizes_batch(const VolArgs* vols, int n, const Workspace& w, int64_t units,
                              int64_t* sizes, hipStream_t st);
hipError_t launch_emit_batch(const Plan& p, int variant, const VolArgs* vols,
                             const int32_t* emap, int64_t gbase, int64_t nblocks, uint32_t* ctr,
                             hipStream_t st);
// Zero n 32-bit words (the work-queue counters) with a kernel.
hipError_t launch_zero_u32(uint32_t* p, int64_t n, hipStream_t st);
// Exact-coordinate vertex welding (weld.hip).
struct WeldArgs {
  int out_kind;          // K_F32 or K_F64: element type of the coordinates
  int64_t nverts, nfaces;
  const void* verts;     // nverts x 3
  void* verts_out;       // welded vertices, room for nverts x 3
  int64_t* faces;        // nfaces x 3, 1-based, renumbered in place
  int64_t* table;        // hash slots: a power of two >= 1.25 nverts
  int64_t nslots;
  uint64_t* bounds;      // room for 63 x-region bounds of a large table
  int64_t* rep;          // nverts: index of the vertex's first appearance
  uint32_t* flag;        // nverts: 1 at first appearances
  int64_t* pos;          // nverts: exclusive prefix of flag
  uint32_t* tilecnt;     // per 4096 vertices: flags (room rounded up to 8 bytes)
  int64_t* tileoff;      // per 4096 vertices: exclusive prefix of tilecnt
  int64_t* total;        // welded count
  int64_t* bad;          // count of face indices outside 1..nverts
};
hipError_t launch_weld(const WeldArgs& a, hipStream_t st);

hipError_t launch_generate(int kind, int dtype, int64_t nx, int64_t ny, int64_t nz,
                           uint64_t seed, double p0, double p1, void* out, hipStream_t st);

// The static device scratch of msh_isosurface_noalloc (noalloc.hip).
constexpr size_t kScratchBytes = MSH_NOALLOC_SCRATCH_BYTES;
// This device's address of the scratch array; call with `device` current.
hipError_t noalloc_setup(int device, void** ptr);
// Once per device, with it current: the lookup-table images of the count
// kernels (lean.hip) and the tetrahedra emit (halo5.hip), built on `st` and
// complete when this returns.
hipError_t tables_setup(int device, hipStream_t st);
hipError_t launch_count_tables(hipStream_t st);
hipError_t launch_mt_tables(hipStream_t st);
// clamp = (0, min(nverts, vcap), 0, min(nfaces, fcap)) from the scan totals,
// and sizes (when not null) = (nverts, nfaces); null totals count as zero.
hipError_t launch_noalloc_sizes(const int64_t* totals, int64_t vcap, int64_t fcap, int64_t* clamp,
                                int64_t* sizes, hipStream_t st);

}  // namespace msh
