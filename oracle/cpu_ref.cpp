// cpu_ref.cpp -- CPU oracle for the isosurface path of Meshing.jl v0.7.0.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference's Julia code
// statement by statement so that tests, the smoke check and the CPU-baseline
// leg of the benchmark can check the HIP library against it.  The product
// (libmeshing_hip.so) never links, loads or calls anything from here.
//
// Followed files (paths are in the reference tree):
//   src/common.jl            _get_cubeindex (:10-20), _no_triangles (:27-29)
//   src/marching_cubes.jl    isosurface MC (:27-64), process_mc_voxel! (:67-92),
//                            vertex_interp (:100-104), mc_vert_points (:111-120)
//   src/marching_tetrahedra.jl tetIx (:11-18), vertId (:29-32), vertPos (:42-55),
//                            getVertId (:67-84), voxEdgeId (:91-95),
//                            procVox (:105-127), isosurface MT (:129-163)
//   src/lut/mc.jl, src/lut/mt.jl  via ref_luts.h (same names, 1-based)
// Julia Base semantics restated here (Base is not part of the reference):
//   LinRange getindex = lerpi: T((1-t)*a + t*b), t = (i-1)/max(1,n-1)
//   exact mixed Float/Int `<`; NaN-propagating min/max with -0 < +0;
//   promotion Int64 < Float32 < Float64, which is exactly C++'s usual
//   arithmetic conversion over {int64_t, float, double}: every Julia
//   expression below is written with the operand types Julia gives it, so
//   the C++ compiler picks the same result type.  Build with
//   -ffp-contract=off (Julia never fuses a*b+c unless asked).
// Naive Surface Nets has no code in v0.7.0; its definition here is this
// project's own (see sn_isosurface below).
//
// Pinning: LinRange's Float64 formula by the exact values of Julia's
// LinRange docstring, which (1-t)*a + t*b reproduces and a + t*(b-a) and
// a + i*step do not; MT's vertex numbering by first appearance in the
// flattened faces and its counts by closed forms over the tetrahedron
// split; the surface-nets spec by its counting and placement invariants.
// Parity unpinned (no Julia run is possible here): LinRange values beyond
// the docstring's, and the noisy-sphere counts of test/runtests.jl:152-172,
// which depend on Julia's MersenneTwister stream.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ref_luts.h"

namespace {

enum Kind { K_INT64 = 0, K_F32 = 1, K_F64 = 2 };

// ---- Julia Base restatements ---------------------------------------------

// `x < y` for a float x and an Int64 y is exact in Julia (base/float.jl).
template <class TF>
bool jl_lt(TF x, int64_t y) {
  TF fy = (TF)y;
  const TF tmax = (TF)INT64_MAX;  // Tf(typemax(Int64)) == 2^63
  return (x < fy) || ((x == fy) && (fy < tmax) && ((int64_t)fy < y));
}
// float vs float: Julia promotes to the wider type, which is exact.
template <class TF>
bool jl_lt(TF x, float y) { return (double)x < (double)y; }
template <class TF>
bool jl_lt(TF x, double y) { return (double)x < y; }

// min/max of IEEE floats: NaN-propagating, -0.0 < +0.0.
template <class C>
C jl_min_same(C x, C y) {
  if (std::isnan(x)) return x;
  if (std::isnan(y)) return y;
  return std::signbit(x - y) ? x : y;
}
template <class C>
C jl_max_same(C x, C y) {
  if (std::isnan(x)) return x;
  if (std::isnan(y)) return y;
  return std::signbit(x - y) ? y : x;
}
template <class A, class B>
auto jl_min(A a, B b) {
  using C = decltype(a + b);
  return jl_min_same<C>((C)a, (C)b);
}
template <class A, class B>
auto jl_max(A a, B b) {
  using C = decltype(a + b);
  return jl_max_same<C>((C)a, (C)b);
}

// LinRange(first, last, n) for one axis.  The element type CX is Float64 for
// Int or Float64 endpoints and Float32 for Float32 endpoints; values are
// held here as double (exact for both) together with the kind.
struct Axis {
  int kind;        // K_F32 or K_F64: the LinRange element type
  double a, b;     // endpoints converted to the element type
  int64_t len, lendiv;
  double operator[](int64_t i) const {  // 1-based
    double t = (double)(i - 1) / (double)lendiv;
    if (kind == K_F32) {
      float af = (float)a, bf = (float)b;
      return (double)(float)((1 - t) * af + t * bf);
    }
    return (1 - t) * a + t * b;
  }
};

Axis make_axis(int src_kind, double first, double last, int64_t n) {
  Axis ax;
  ax.kind = (src_kind == K_F32) ? K_F32 : K_F64;
  if (ax.kind == K_F32) {
    ax.a = (double)(float)first;
    ax.b = (double)(float)last;
  } else {
    ax.a = first;
    ax.b = last;
  }
  if (n < 0) throw std::invalid_argument("LinRange: negative length");
  if (n == 1 && !(ax.a == ax.b))
    throw std::invalid_argument("LinRange: endpoints differ for length 1");
  ax.len = n;
  ax.lendiv = n - 1 > 1 ? n - 1 : 1;
  return ax;
}

// One axis of p1 .+ mu .* (p2 .- p1), with p1, p2 of the axis's element
// type CX.  Returns the value of type promote(M, CX) widened to double.
template <class M>
double interp_axis(int cx, M mu, double p1, double p2) {
  if (cx == K_F32) {
    float a = (float)p1, b = (float)p2;
    return (double)(a + mu * (b - a));
  }
  return (double)(p1 + mu * (p2 - p1));
}

// ---- common.jl -------------------------------------------------------------

template <class T, class I>
uint8_t get_cubeindex(const T* v, I iso) {
  uint8_t c = jl_lt(v[0], iso) ? 0x01 : 0x00;
  if (jl_lt(v[1], iso)) c |= 0x02;
  if (jl_lt(v[2], iso)) c |= 0x04;
  if (jl_lt(v[3], iso)) c |= 0x08;
  if (jl_lt(v[4], iso)) c |= 0x10;
  if (jl_lt(v[5], iso)) c |= 0x20;
  if (jl_lt(v[6], iso)) c |= 0x40;
  if (jl_lt(v[7], iso)) c |= 0x80;
  return c;
}
inline bool no_triangles(uint8_t c) { return c == 0x00 || c == 0xff; }

// sdf[i, j, k], 1-based.  The data may be an x-slab of a larger array: it
// then holds planes xoff+1 .. xoff+sx of the array (sx x-samples per row),
// while nx stays the array's size along x.
template <class T>
struct Grid {
  const T* p;
  int64_t nx, ny, nz;
  int64_t xoff = 0, sx = 0;  // sx == 0: the whole array (sx = nx)
  T operator()(int64_t i, int64_t j, int64_t k) const {
    const int64_t w = sx ? sx : nx;
    return p[(i - 1 - xoff) + w * ((j - 1) + ny * (k - 1))];
  }
};

template <class Out>
struct Result {
  std::vector<Out> v;        // 3 per vertex
  std::vector<int64_t> f;    // 3 per face, 1-based
};

// ---- marching_cubes.jl -------------------------------------------------------

struct Pt { double c[3]; };

// Cells with xi in [xlo, xhi] (default: all).  Face indices count from the
// first vertex this call emits, so a slab's result is the serial result's
// slice with its faces shifted by the vertices of the slabs before it.
template <class T, class I, class Out>
void mc_isosurface(const Grid<T>& sdf, I iso, const Axis (&ax)[3], Result<Out>& r,
                   int64_t xlo = 1, int64_t xhi = -1) {
  const int64_t nx = sdf.nx, ny = sdf.ny, nz = sdf.nz;
  const Axis &xp = ax[0], &yp = ax[1], &zp = ax[2];
  if (xhi < 0) xhi = nx - 1;
  for (int64_t xi = xlo; xi <= xhi; ++xi)
    for (int64_t yi = 1; yi <= ny - 1; ++yi)
      for (int64_t zi = 1; zi <= nz - 1; ++zi) {
        const T iso_vals[8] = {sdf(xi, yi, zi),         sdf(xi + 1, yi, zi),
                               sdf(xi + 1, yi + 1, zi), sdf(xi, yi + 1, zi),
                               sdf(xi, yi, zi + 1),     sdf(xi + 1, yi, zi + 1),
                               sdf(xi + 1, yi + 1, zi + 1), sdf(xi, yi + 1, zi + 1)};
        const uint8_t cubeindex = get_cubeindex(iso_vals, iso);
        if (no_triangles(cubeindex)) continue;
        // mc_vert_points
        const Pt points[8] = {{{xp[xi], yp[yi], zp[zi]}},
                              {{xp[xi + 1], yp[yi], zp[zi]}},
                              {{xp[xi + 1], yp[yi + 1], zp[zi]}},
                              {{xp[xi], yp[yi + 1], zp[zi]}},
                              {{xp[xi], yp[yi], zp[zi + 1]}},
                              {{xp[xi + 1], yp[yi], zp[zi + 1]}},
                              {{xp[xi + 1], yp[yi + 1], zp[zi + 1]}},
                              {{xp[xi], yp[yi + 1], zp[zi + 1]}}};
        // process_mc_voxel!
        const int64_t fct = (int64_t)(r.v.size() / 3);
        const uint8_t* vert_to_add = _mc_verts[cubeindex - 1];
        for (int i = 1; i <= 12; ++i) {
          const int vt = vert_to_add[i - 1];
          if (vt == 0) break;
          const uint8_t* ed = _mc_edge_list[vt - 1];
          const T valp1 = iso_vals[ed[0] - 1], valp2 = iso_vals[ed[1] - 1];
          // vertex_interp: mu = (iso - valp1) / (valp2 - valp1)
          const auto mu = (iso - valp1) / (valp2 - valp1);
          for (int k = 0; k < 3; ++k)
            r.v.push_back((Out)interp_axis(ax[k].kind, mu, points[ed[0] - 1].c[k],
                                           points[ed[1] - 1].c[k]));
        }
        const uint8_t* offsets = _mc_connectivity[_mc_eq_mapping[cubeindex - 1] - 1];
        r.f.push_back(fct + 3);
        r.f.push_back(fct + 2);
        r.f.push_back(fct + 1);
        for (int i : {1, 4, 7, 10}) {
          if (offsets[i - 1] == 0) break;
          r.f.push_back(fct + offsets[i + 1]);
          r.f.push_back(fct + offsets[i]);
          r.f.push_back(fct + offsets[i - 1]);
        }
      }
}

// ---- marching_tetrahedra.jl ------------------------------------------------

inline uint8_t tetIx(int tIx, uint8_t cubeindex) {
  const uint8_t v1 = subTetsMask[tIx - 1][0];
  const uint8_t v2 = subTetsMask[tIx - 1][1];
  uint8_t ix = (0x01 & cubeindex) | ((0x40 & cubeindex) >> 0x03);
  if ((v1 & cubeindex) != 0) ix |= 0x02;
  if ((v2 & cubeindex) != 0) ix |= 0x04;
  return ix;
}

inline int64_t vertId(int e, int64_t x, int64_t y, int64_t z, int64_t nx, int64_t ny) {
  const int64_t* d = voxCrnrPos[voxEdgeCrnrs[e - 1][0] - 1];
  return (int64_t)voxEdgeDir[e - 1] + 7 * (x - 1 + d[0] + nx * (y - 1 + d[1] + ny * (z - 1 + d[2])));
}

inline int voxEdgeId(int subTetIx, int tetEdgeIx) {
  const int src = subTets[subTetIx - 1][tetEdgeCrnrs[tetEdgeIx - 1][0] - 1];
  const int tgt = subTets[subTetIx - 1][tetEdgeCrnrs[tetEdgeIx - 1][1] - 1];
  return voxEdgeIx[src - 1][tgt - 1];
}

template <class T, class I, class E, class Out>
struct MT {
  const Grid<T>& sdf;
  I iso;
  E eps;
  const Axis (&ax)[3];
  Result<Out>& r;
  std::unordered_map<int64_t, int64_t> vts;

  MT(const Grid<T>& s, I i, E e, const Axis (&a)[3], Result<Out>& res)
      : sdf(s), iso(i), eps(e), ax(a), r(res) {}

  void vertPos(int e, int64_t x, int64_t y, int64_t z, const T* vals) {
    const uint8_t* ixs = voxEdgeCrnrs[e - 1];
    const T srcVal = vals[ixs[0] - 1];
    const T tgtVal = vals[ixs[1] - 1];
    const auto a = jl_min(jl_max((iso - srcVal) / (tgtVal - srcVal), eps), T(1) - eps);
    const auto b = T(1) - a;
    const int64_t* c1 = voxCrnrPos[ixs[0] - 1];
    const int64_t* c2 = voxCrnrPos[ixs[1] - 1];
    const int64_t base[3] = {x, y, z};
    for (int k = 0; k < 3; ++k) {
      const auto w = c1[k] * b + c2[k] * a;
      double res;
      if (ax[k].kind == K_F32) {
        const float lo = (float)ax[k][base[k]], hi = (float)ax[k][base[k] + 1];
        const float d = hi - lo;
        res = (double)(lo + w * d);
      } else {
        const double lo = ax[k][base[k]], hi = ax[k][base[k] + 1];
        const double d = hi - lo;
        res = (double)(lo + w * d);
      }
      r.v.push_back((Out)res);
    }
  }

  int64_t getVertId(int e, int64_t x, int64_t y, int64_t z, const T* vals) {
    const int64_t vId = vertId(e, x, y, z, sdf.nx, sdf.ny);
    auto it = vts.find(vId);
    if (it != vts.end()) return it->second;
    vertPos(e, x, y, z, vals);
    const int64_t n = (int64_t)(r.v.size() / 3);
    vts[vId] = n;
    return n;
  }

  void procVox(const T* vals, int64_t x, int64_t y, int64_t z, uint8_t cubeindex) {
    for (int i = 1; i <= 6; ++i) {
      const uint8_t tIx = tetIx(i, cubeindex);
      if (tIx == 0x00 || tIx == 0x0f) continue;
      const uint8_t* e = tetTri[tIx - 1];
      // arguments are evaluated left to right in Julia
      const int64_t a = getVertId(voxEdgeId(i, e[0]), x, y, z, vals);
      const int64_t b = getVertId(voxEdgeId(i, e[1]), x, y, z, vals);
      const int64_t c = getVertId(voxEdgeId(i, e[2]), x, y, z, vals);
      r.f.push_back(a);
      r.f.push_back(b);
      r.f.push_back(c);
      if (e[3] == 0) continue;
      const int64_t d = getVertId(voxEdgeId(i, e[3]), x, y, z, vals);
      const int64_t f = getVertId(voxEdgeId(i, e[4]), x, y, z, vals);
      const int64_t g = getVertId(voxEdgeId(i, e[5]), x, y, z, vals);
      r.f.push_back(d);
      r.f.push_back(f);
      r.f.push_back(g);
    }
  }

  void run() {
    const int64_t nx = sdf.nx, ny = sdf.ny, nz = sdf.nz;
    for (int64_t i = 1; i <= nx - 1; ++i)
      for (int64_t j = 1; j <= ny - 1; ++j)
        for (int64_t k = 1; k <= nz - 1; ++k) {
          const T vals[8] = {sdf(i, j, k),         sdf(i, j + 1, k),
                             sdf(i + 1, j + 1, k), sdf(i + 1, j, k),
                             sdf(i, j, k + 1),     sdf(i, j + 1, k + 1),
                             sdf(i + 1, j + 1, k + 1), sdf(i + 1, j, k + 1)};
          const uint8_t cubeindex = get_cubeindex(vals, iso);
          if (no_triangles(cubeindex)) continue;
          procVox(vals, i, j, k, cubeindex);
        }
  }
};

// ---- Naive Surface Nets (this project's own definition) -------------------
//
// Meshing.jl v0.7.0 has no Surface Nets code.  Definition used, after the
// published method (0fps.net, "Smooth voxel terrain, part 2"), in the
// reference's conventions:
//  * cells, corner order, classification and skip rule as Marching Cubes;
//  * one vertex per active cell, numbered by active-cell order in the MC
//    loop order (x outermost, z innermost), 1-based;
//  * its position is the mean over the cell's crossing edges, in MC edge
//    order 1..12, of vertex_interp's point: per axis the points are summed
//    in promote(M, CX) starting from zero, divided by the count in that
//    type, and rounded to float(FT) like MC;
//  * for each axis a in (x, y, z), the grid edge leaving the cell's first
//    corner along a is "crossing" when its two corner bits differ.  If the
//    cell index along both other axes u=(a+1)%3, w=(a+2)%3 is at least 2,
//    the four cells around that edge -- c, c-u, c-u-w, c-w -- form a quad,
//    in that order when the first corner is below iso and reversed
//    (c, c-w, c-u-w, c-u) otherwise, split as (q0,q1,q2), (q0,q2,q3).
template <class T, class I, class Out>
void sn_isosurface(const Grid<T>& sdf, I iso, const Axis (&ax)[3], Result<Out>& r) {
  const int64_t nx = sdf.nx, ny = sdf.ny, nz = sdf.nz;
  const int64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
  if (cx <= 0 || cy <= 0 || cz <= 0) return;
  std::vector<int64_t> id((size_t)(cx * cy * cz), 0);
  auto cell = [&](int64_t i, int64_t j, int64_t k) -> int64_t& {
    return id[(size_t)(((i - 1) * cy + (j - 1)) * cz + (k - 1))];
  };
  static const int corner_of_axis[3] = {2, 4, 5};  // MC corners (1,0,0),(0,1,0),(0,0,1)
  const Axis &xp = ax[0], &yp = ax[1], &zp = ax[2];
  for (int64_t xi = 1; xi <= cx; ++xi)
    for (int64_t yi = 1; yi <= cy; ++yi)
      for (int64_t zi = 1; zi <= cz; ++zi) {
        const T v[8] = {sdf(xi, yi, zi),         sdf(xi + 1, yi, zi),
                        sdf(xi + 1, yi + 1, zi), sdf(xi, yi + 1, zi),
                        sdf(xi, yi, zi + 1),     sdf(xi + 1, yi, zi + 1),
                        sdf(xi + 1, yi + 1, zi + 1), sdf(xi, yi + 1, zi + 1)};
        const uint8_t c = get_cubeindex(v, iso);
        if (no_triangles(c)) continue;
        const Pt p[8] = {{{xp[xi], yp[yi], zp[zi]}},
                         {{xp[xi + 1], yp[yi], zp[zi]}},
                         {{xp[xi + 1], yp[yi + 1], zp[zi]}},
                         {{xp[xi], yp[yi + 1], zp[zi]}},
                         {{xp[xi], yp[yi], zp[zi + 1]}},
                         {{xp[xi + 1], yp[yi], zp[zi + 1]}},
                         {{xp[xi + 1], yp[yi + 1], zp[zi + 1]}},
                         {{xp[xi], yp[yi + 1], zp[zi + 1]}}};
        // accumulate in promote(M, CX) per axis
        using M = decltype((iso - v[0]) / (v[1] - v[0]));
        float accf[3] = {0.f, 0.f, 0.f};
        double accd[3] = {0.0, 0.0, 0.0};
        bool is_f[3];
        for (int k = 0; k < 3; ++k)
          is_f[k] = (sizeof(M) == 4) && ax[k].kind == K_F32;
        int n = 0;
        for (int e = 1; e <= 12; ++e) {
          const int a = _mc_edge_list[e - 1][0], b = _mc_edge_list[e - 1][1];
          const bool ba = (c >> (a - 1)) & 1, bb = (c >> (b - 1)) & 1;
          if (ba == bb) continue;
          ++n;
          const M mu = (iso - v[a - 1]) / (v[b - 1] - v[a - 1]);
          for (int k = 0; k < 3; ++k) {
            const double q = interp_axis(ax[k].kind, mu, p[a - 1].c[k], p[b - 1].c[k]);
            if (is_f[k]) accf[k] = accf[k] + (float)q;
            else accd[k] = accd[k] + q;
          }
        }
        for (int k = 0; k < 3; ++k) {
          const double q = is_f[k] ? (double)(accf[k] / (float)n) : accd[k] / (double)n;
          r.v.push_back((Out)q);
        }
        const int64_t me = (int64_t)(r.v.size() / 3);
        cell(xi, yi, zi) = me;
        const int64_t idx[3] = {xi, yi, zi};
        const bool below0 = c & 1;
        for (int a = 0; a < 3; ++a) {
          const bool bf = (c >> (corner_of_axis[a] - 1)) & 1;
          if (bf == below0) continue;
          const int u = (a + 1) % 3, w = (a + 2) % 3;
          if (idx[u] < 2 || idx[w] < 2) continue;
          int64_t cu[3] = {xi, yi, zi}, cw[3] = {xi, yi, zi}, cuw[3] = {xi, yi, zi};
          cu[u] -= 1;
          cw[w] -= 1;
          cuw[u] -= 1;
          cuw[w] -= 1;
          const int64_t q1 = cell(cu[0], cu[1], cu[2]);
          const int64_t q2 = cell(cuw[0], cuw[1], cuw[2]);
          const int64_t q3 = cell(cw[0], cw[1], cw[2]);
          int64_t quad[4];
          if (below0) {
            quad[0] = me; quad[1] = q1; quad[2] = q2; quad[3] = q3;
          } else {
            quad[0] = me; quad[1] = q3; quad[2] = q2; quad[3] = q1;
          }
          r.f.push_back(quad[0]); r.f.push_back(quad[1]); r.f.push_back(quad[2]);
          r.f.push_back(quad[0]); r.f.push_back(quad[2]); r.f.push_back(quad[3]);
        }
      }
}

// ---- dispatch over the Julia types --------------------------------------------

struct Scalar {
  int kind;
  double f;
  int64_t i;
};

int out_kind_of(int dtype, const int* kinds, const Scalar& iso, const Scalar* eps) {
  // float(promote_type(eltype(first(X)), .., eltype(T), typeof(iso)[, typeof(eps)]))
  bool f64 = dtype == K_F64 || iso.kind == K_F64 || kinds[0] == K_F64 ||
             kinds[1] == K_F64 || kinds[2] == K_F64 || (eps && eps->kind == K_F64);
  return f64 ? K_F64 : K_F32;
}

int g_threads = 1;  // > 1: Marching Cubes over x-slabs on that many threads
// Marching Cubes over the cells with xi in [g_xlo, g_xhi] only (g_xhi < 0:
// all); set by oracle_isosurface_xslab for one call.
int64_t g_xlo = 1, g_xhi = -1;

// The multi-threaded form of mc_isosurface (a CPU baseline over all cores,
// not a restatement): x-slabs on separate threads, concatenated in x order
// with each slab's faces shifted by the vertices before it, which is the
// serial result exactly because a cube's faces name only its own vertices.
template <class T, class I, class Out>
void mc_isosurface_threads(const Grid<T>& g, I iso, const Axis (&ax)[3], Result<Out>& r,
                           int nt) {
  const int64_t ncx = g.nx - 1;
  if (ncx <= 0) return;
  if (nt > ncx) nt = (int)ncx;
  std::vector<Result<Out>> part(nt);
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      const int64_t lo = 1 + ncx * t / nt, hi = ncx * (t + 1) / nt;
      mc_isosurface<T, I, Out>(g, iso, ax, part[t], lo, hi);
    });
  for (auto& x : th) x.join();
  size_t nv = 0, nf = 0;
  for (auto& p : part) nv += p.v.size(), nf += p.f.size();
  r.v.reserve(nv);
  r.f.reserve(nf);
  for (auto& p : part) {
    const int64_t shift = (int64_t)(r.v.size() / 3);
    r.v.insert(r.v.end(), p.v.begin(), p.v.end());
    for (int64_t i : p.f) r.f.push_back(i + shift);
    Result<Out>().v.swap(p.v);
  }
}

template <class T, class I, class E, class Out>
void run_method(int method, const Grid<T>& g, I iso, E eps, const Axis (&ax)[3], Result<Out>& r) {
  if (method == 0 && g_threads > 1) mc_isosurface_threads<T, I, Out>(g, iso, ax, r, g_threads);
  else if (method == 0) mc_isosurface<T, I, Out>(g, iso, ax, r, g_xlo, g_xhi);
  else if (method == 1) MT<T, I, E, Out>(g, iso, eps, ax, r).run();
  else sn_isosurface<T, I, Out>(g, iso, ax, r);
}

template <class T, class I, class E>
void run_out(int method, int out_kind, const Grid<T>& g, I iso, E eps, const Axis (&ax)[3],
             void** verts, int64_t* nv, int64_t** faces, int64_t* nf) {
  auto finish = [&](auto& r) {
    using O = typename std::decay_t<decltype(r.v)>::value_type;
    *nv = (int64_t)(r.v.size() / 3);
    *nf = (int64_t)(r.f.size() / 3);
    *verts = std::malloc(r.v.size() * sizeof(O) + 1);
    *faces = (int64_t*)std::malloc(r.f.size() * sizeof(int64_t) + 1);
    if (!r.v.empty()) std::memcpy(*verts, r.v.data(), r.v.size() * sizeof(O));
    if (!r.f.empty()) std::memcpy(*faces, r.f.data(), r.f.size() * sizeof(int64_t));
  };
  if (out_kind == K_F32) {
    Result<float> r;
    run_method<T, I, E, float>(method, g, iso, eps, ax, r);
    finish(r);
  } else {
    Result<double> r;
    run_method<T, I, E, double>(method, g, iso, eps, ax, r);
    finish(r);
  }
}

template <class T, class I>
void run_eps(int method, int out_kind, const Grid<T>& g, I iso, const Scalar& eps,
             const Axis (&ax)[3], void** v, int64_t* nv, int64_t** f, int64_t* nf) {
  if (method != 1 || eps.kind == K_F64) run_out<T, I, double>(method, out_kind, g, iso, eps.f, ax, v, nv, f, nf);
  else if (eps.kind == K_F32) run_out<T, I, float>(method, out_kind, g, iso, (float)eps.f, ax, v, nv, f, nf);
  else run_out<T, I, int64_t>(method, out_kind, g, iso, eps.i, ax, v, nv, f, nf);
}

template <class T>
void run_iso(int method, int out_kind, const Grid<T>& g, const Scalar& iso, const Scalar& eps,
             const Axis (&ax)[3], void** v, int64_t* nv, int64_t** f, int64_t* nf) {
  if (iso.kind == K_F64) run_eps<T, double>(method, out_kind, g, iso.f, eps, ax, v, nv, f, nf);
  else if (iso.kind == K_F32) run_eps<T, float>(method, out_kind, g, (float)iso.f, eps, ax, v, nv, f, nf);
  else run_eps<T, int64_t>(method, out_kind, g, iso.i, eps, ax, v, nv, f, nf);
}

thread_local std::string g_err;

}  // namespace

extern "C" {

// Full isosurface.  method 0 MC, 1 MT, 2 SN.  dtype K_F32/K_F64 (the array's
// eltype); iso/eps given as (kind, double value, int64 value).  Axis kinds
// are the eltypes of first(X), first(Y), first(Z).  Outputs are malloc'ed
// (free with oracle_free); *out_kind is float(FT).  Returns 0, or -1 with
// oracle_last_error() set (e.g. LinRange length-1 with differing endpoints).
int oracle_isosurface(int method, int dtype, const void* sdf, int64_t nx, int64_t ny, int64_t nz,
                      int iso_kind, double iso_f, int64_t iso_i, int eps_kind, double eps_f,
                      int64_t eps_i, const int* kinds, const double* lo, const double* hi,
                      void** verts, int64_t* nverts, int64_t** faces, int64_t* nfaces,
                      int* out_kind) {
  try {
    Axis ax[3] = {make_axis(kinds[0], lo[0], hi[0], nx), make_axis(kinds[1], lo[1], hi[1], ny),
                  make_axis(kinds[2], lo[2], hi[2], nz)};
    Scalar iso{iso_kind, iso_f, iso_i}, eps{eps_kind, eps_f, eps_i};
    *out_kind = out_kind_of(dtype, kinds, iso, method == 1 ? &eps : nullptr);
    if (dtype == K_F32) {
      Grid<float> g{(const float*)sdf, nx, ny, nz};
      run_iso<float>(method, *out_kind, g, iso, eps, ax, verts, nverts, faces, nfaces);
    } else if (dtype == K_F64) {
      Grid<double> g{(const double*)sdf, nx, ny, nz};
      run_iso<double>(method, *out_kind, g, iso, eps, ax, verts, nverts, faces, nfaces);
    } else {
      throw std::invalid_argument("unsupported field element type");
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Marching Cubes over the cells xi = xlo..xhi (1-based) of an nx x ny x nz
// array of which only an x-slab is given: `slab` holds x-planes
// xoff+1 .. xoff+sx (sx samples per row, x fastest).  Coordinates are those
// of the whole array (LinRange over nx).  The result is the slice of the
// whole array's result that these cells emit, with face indices counted
// from the slice's first vertex (the loop is x-outermost,
// src/marching_cubes.jl:40, and a cube's faces name only its own vertices).
int oracle_isosurface_xslab(int dtype, const void* slab, int64_t nx, int64_t ny, int64_t nz,
                            int64_t xoff, int64_t sx, int64_t xlo, int64_t xhi, int iso_kind,
                            double iso_f, int64_t iso_i, const int* kinds, const double* lo,
                            const double* hi, void** verts, int64_t* nverts, int64_t** faces,
                            int64_t* nfaces, int* out_kind) {
  try {
    if (xlo < 1 || xhi > nx - 1 || xlo > xhi || xlo < xoff + 1 || xhi + 1 > xoff + sx ||
        sx < 1 || xoff + sx > nx)
      throw std::invalid_argument("x-slab does not hold the requested cells");
    Axis ax[3] = {make_axis(kinds[0], lo[0], hi[0], nx), make_axis(kinds[1], lo[1], hi[1], ny),
                  make_axis(kinds[2], lo[2], hi[2], nz)};
    Scalar iso{iso_kind, iso_f, iso_i}, eps{K_F64, 1e-3, 0};
    *out_kind = out_kind_of(dtype, kinds, iso, nullptr);
    const int saved = g_threads;
    g_threads = 1;
    g_xlo = xlo;
    g_xhi = xhi;
    try {
      if (dtype == K_F32) {
        Grid<float> g{(const float*)slab, nx, ny, nz, xoff, sx};
        run_iso<float>(0, *out_kind, g, iso, eps, ax, verts, nverts, faces, nfaces);
      } else if (dtype == K_F64) {
        Grid<double> g{(const double*)slab, nx, ny, nz, xoff, sx};
        run_iso<double>(0, *out_kind, g, iso, eps, ax, verts, nverts, faces, nfaces);
      } else {
        throw std::invalid_argument("unsupported field element type");
      }
    } catch (...) {
      g_threads = saved;
      g_xlo = 1;
      g_xhi = -1;
      throw;
    }
    g_threads = saved;
    g_xlo = 1;
    g_xhi = -1;
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Threads used by later oracle_isosurface calls for Marching Cubes (1:
// the single-threaded restatement; MT and SN always run on one thread).
// A process-wide setting of this test library, not for concurrent use.
void oracle_set_threads(int n) { g_threads = n > 1 ? n : 1; }

void oracle_free(void* p) { std::free(p); }
const char* oracle_last_error(void) { return g_err.c_str(); }

// LinRange(first, last, n)[i] for i = 1..n, written to out (double).
int oracle_linrange(int kind, double first, double last, int64_t n, double* out) {
  try {
    Axis a = make_axis(kind, first, last, n);
    for (int64_t i = 1; i <= n; ++i) out[i - 1] = a[i];
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// _get_cubeindex over 8 Float64 values with a Float64 iso.
int oracle_cubeindex_f64(const double* v, double iso) { return get_cubeindex(v, iso); }

// vertex_interp(iso, p1, p2, v1, v2) with Int64 arguments promoted as in
// the reference's test (Int iso, Int points, Int values -> Float64 result).
void oracle_vertex_interp_i64(int64_t iso, const int64_t* p1, const int64_t* p2, int64_t v1,
                              int64_t v2, double* out) {
  const double mu = (double)(iso - v1) / (double)(v2 - v1);  // Int / Int -> Float64
  for (int k = 0; k < 3; ++k) out[k] = (double)p1[k] + mu * (double)(p2[k] - p1[k]);
}

}  // extern "C"
