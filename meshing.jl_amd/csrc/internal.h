// internal.h -- shared between the C ABI (capi.cpp) and the kernels
// (kernels.hip).  Not installed; the public interface is include/meshing_hip.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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
};

struct Workspace {
  uint32_t* cnt;      // per unit: vertices | faces << 16
  int64_t* voff;      // per unit: exclusive prefix of vertices
  int64_t* foff;      // per unit: exclusive prefix of faces
  uint64_t* status;   // per scan tile: look-back flag + aggregates
  int64_t* incl;      // per scan tile: inclusive prefixes (v, f)
  uint32_t* counter;  // scan tile ticket
  int64_t* totals;    // (nverts, nfaces)
  double* coords;     // xp[nx], yp[ny], zp[nz] as double
};

constexpr int kScanItems = 8;
constexpr int kScanTile = 256 * kScanItems;

inline int64_t scan_tiles(int64_t units) { return (units + kScanTile - 1) / kScanTile; }

hipError_t launch_coords(const Plan& p, const Geom& g, double* coords, hipStream_t st);
hipError_t launch_count(const Plan& p, const Geom& g, const void* field, uint32_t* cnt,
                        hipStream_t st);
hipError_t launch_scan(const Geom& g, const Workspace& w, hipStream_t st);
hipError_t launch_emit(const Plan& p, const Geom& g, const void* field, const Workspace& w,
                       void* verts, int64_t* faces, hipStream_t st);
hipError_t launch_generate(int kind, int dtype, int64_t nx, int64_t ny, int64_t nz,
                           uint64_t seed, double p0, double p1, void* out, hipStream_t st);

}  // namespace msh
