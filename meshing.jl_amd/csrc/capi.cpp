// capi.cpp -- the C ABI of libmeshing_hip.so (declared in include/meshing_hip.h).
//
// Replaces the reference's `isosurface(sdf, method, X, Y, Z)` specialisations
// (src/isosurface.jl:30-32, src/marching_cubes.jl:27-64,
// src/marching_tetrahedra.jl:129-163) with a two-phase call: msh_count
// (classify + count + scan, returns exact sizes) and msh_extract (emit into
// caller-owned buffers).  Julia's type promotion is resolved here, on the
// host, into the handful of arithmetic types the kernels are instantiated
// for (internal.h, struct Plan).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/meshing_hip.h"
#include "internal.h"

using namespace msh;

namespace {

// One kernel launch of a batch: the volumes that share its instantiation.
struct Launch {
  Plan plan;                // types (and method) of the group
  int variant = 0;          // emit variant (emit launches only)
  int64_t gbase = 0;        // first block in the block -> volume map
  int64_t nblocks = 0;
  double cost = 0;          // estimated work: cells weighted by method
};

// Estimated work of a volume in a batched launch, for spreading the launches
// over streams: its tile-cells (a partly filled tile costs about what a full
// one does), and for an emit weighted by method (MC 1, surface nets 2,
// tetrahedra 3: the 256^3 gyroid emits take 0.056, 0.12 and 0.17 ms);
// counting costs the same per cell for every method.
double cell_cost(const VolArgs& v, const Plan& p, bool emit) {
  const double zc = (double)(v.g.zseg > 0 ? v.g.zseg : 0);
  return emit ? (double)v.ntiles * zc * (p.method == MT ? 3.0 : p.method == SN ? 2.0 : 1.0)
              : (double)v.cntiles * zc;
}

// Launch order for spreading: most work first (each goes to the stream with
// the least work so far, so the short launches fill in at the end).
std::vector<size_t> by_cost(const std::vector<Launch>& ls) {
  std::vector<size_t> o(ls.size());
  for (size_t i = 0; i < o.size(); ++i) o[i] = i;
  std::stable_sort(o.begin(), o.end(), [&](size_t x, size_t y) { return ls[x].cost > ls[y].cost; });
  return o;
}

// A captured phase of the batched form.  A batch whose arguments and
// device buffers repeat those of the previous call runs its device work from
// a hipGraph: the second such call captures it, later ones replay it.  The
// key is the bytes that decide every launch and copy of the phase.
struct PhaseGraph {
  std::vector<char> last;   // key of the previous call of this phase
  std::vector<char> key;    // key the graph was captured for
  hipGraphExec_t exec = nullptr;
  int64_t replays = 0;
};

// State of the last successful msh_batch_count.
struct BatchState {
  bool have = false;
  std::vector<msh_field> fields;
  std::vector<msh_params> params;
  std::vector<Plan> plans;
  std::vector<VolArgs> vols;       // host copy of the device descriptors
  std::vector<Launch> counts, emits;
  std::vector<int64_t> sizes;      // 4 per volume: first vertex, nverts, first face, nfaces
  int64_t units = 0, ncoords = 0, cblocks = 0, eblocks = 0;
  size_t ctr_off = 0;              // work-queue counters in the batch workspace: 8 per
                                   // count launch, then 8 per emit launch
};

constexpr int kSpreadWays = 3;  // streams a batched phase's launches spread over

struct Ctx {
  int device = 0;
  hipStream_t own = nullptr;
  hipStream_t stream = nullptr;
  std::string err;
  // grow-only device buffers
  void* ws = nullptr;
  size_t ws_bytes = 0;
  void* fstage = nullptr;
  size_t fstage_bytes = 0;
  void* ostage = nullptr;
  size_t ostage_bytes = 0;
  void* bws = nullptr;        // batch: descriptors, block maps, sizes
  size_t bws_bytes = 0;
  void* hbatch = nullptr;     // batch: pinned host copy of descriptors and sizes
  size_t hbatch_bytes = 0;
  hipEvent_t upload = nullptr;  // last descriptor upload from hbatch
  bool upload_pending = false;
  // batch: more streams (default priority) that a phase's launches spread
  // over, forked from and joined back to `stream` by events
  hipStream_t aux[kSpreadWays - 1] = {};
  hipEvent_t fork = nullptr, join[kSpreadWays - 1] = {};
  int64_t* pinned = nullptr;  // 4 x int64 for the totals
  void* scratch = nullptr;    // this device's static scratch (msh_isosurface_noalloc)
  void* wws = nullptr;        // weld: hash table, scan workspace, indices
  size_t wws_bytes = 0;
  // emit kernel family per (method, element type, size class), -1 = default
  // order; set by msh_autotune or msh_set_kernel_family
  int family[3][3][kSizeClasses];
  Ctx() {
    for (auto& m : family)
      for (auto& t : m)
        for (int& f : t) f = -1;
  }
  // state of the last successful msh_count
  bool have = false;
  msh_field field{};
  msh_params params{};
  Plan plan{};
  Geom geom{};
  Workspace w{};
  VolArgs vol{};
  int64_t nverts = 0, nfaces = 0;
  BatchState batch;
  PhaseGraph gcount, gemit;  // batched count and emit phases
};

void graph_drop(PhaseGraph* g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  g->exec = nullptr;
  g->key.clear();
}

template <class T>
void key_add(std::vector<char>* k, const T* p, size_t n) {
  const char* b = (const char*)p;
  k->insert(k->end(), b, b + sizeof(T) * n);
}

// Run `enqueue` (device work on c->stream) eagerly, or through the phase's
// graph: replayed when `key` matches the captured one, captured when it
// matches the previous call's (a batch that repeats).  `graphable` is false
// for work a graph must not hold (pageable host copies, the null stream).
template <class F>
hipError_t run_phase(Ctx* c, PhaseGraph* g, std::vector<char> key, bool graphable, F enqueue) {
  hipError_t e = hipSuccess;
  const bool repeat = graphable && key == g->last;
  g->last = key;
  if (!repeat) return enqueue();
  if (!g->exec || g->key != key) {
    graph_drop(g);
    hipGraph_t graph = nullptr;
    e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return enqueue();  // not capturable here: run eagerly
    const hipError_t eq = enqueue();
    e = hipStreamEndCapture(c->stream, &graph);
    if (eq != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return eq;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {  // nothing ran: run the phase eagerly
      g->exec = nullptr;
      (void)hipGetLastError();
      return enqueue();
    }
    g->key = key;
  }
  ++g->replays;
  return hipGraphLaunch(g->exec, c->stream);
}

// A phase's launches spread over up to kSpreadWays streams: the context's
// and more, forked after the work queued so far and joined before anything
// queued later.  Persistent launches of different kernels then run side by
// side, so the CUs one launch leaves as its last tiles drain take the next
// launch's tiles instead of idling; in a captured phase the streams become
// parallel branches of the graph.  The launches share nothing but the
// inputs they read.
struct Spread {
  Ctx* c;
  int ways = 1;    // streams launches go to
  int forked = 0;  // aux streams that waited on the fork
  std::vector<double> load;
  hipError_t start(int nlaunch) {
    ways = nlaunch < kSpreadWays ? (nlaunch < 1 ? 1 : nlaunch) : kSpreadWays;
    load.assign((size_t)ways, 0);
    hipError_t e = hipSuccess;
    for (int i = 0; i < ways - 1 && e == hipSuccess; ++i) {
      if (!c->aux[i]) e = hipStreamCreateWithFlags(&c->aux[i], hipStreamNonBlocking);
      if (e == hipSuccess && !c->join[i])
        e = hipEventCreateWithFlags(&c->join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess && ways > 1 && !c->fork)
      e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    if (e == hipSuccess && ways > 1) e = hipEventRecord(c->fork, c->stream);
    for (int i = 0; i < ways - 1 && e == hipSuccess; ++i) {
      e = hipStreamWaitEvent(c->aux[i], c->fork, 0);
      if (e == hipSuccess) forked = i + 1;
    }
    if (e != hipSuccess) ways = 1 + forked;  // launch only on streams that forked
    return e;
  }
  // the stream with the least work so far takes a launch of `work`
  hipStream_t next(double work) {
    int k = 0;
    for (int i = 1; i < ways; ++i)
      if (load[i] < load[k]) k = i;
    load[k] += work;
    return k == 0 ? c->stream : c->aux[k - 1];
  }
  // joins every stream the fork reached, also after a failed launch, so a
  // capture in progress always ends joined
  hipError_t finish() {
    hipError_t first = hipSuccess;
    for (int i = 0; i < forked; ++i) {
      hipError_t e = hipEventRecord(c->join[i], c->aux[i]);
      if (e == hipSuccess) e = hipStreamWaitEvent(c->stream, c->join[i], 0);
      if (first == hipSuccess) first = e;
    }
    forked = 0;
    return first;
  }
};

int fail(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int fail(std::string* err, int code, const std::string& msg) {
  if (err) *err = msg;
  return code;
}

int hip_fail(Ctx* c, hipError_t e, const char* what) {
  return fail(c, e == hipErrorOutOfMemory ? MSH_ERR_OUT_OF_MEMORY : MSH_ERR_DEVICE,
              std::string(what) + ": " + hipGetErrorString(e));
}

bool valid_kind(int k) { return k == K_INT64 || k == K_F32 || k == K_F64; }

// Smallest float >= the exact value of x (an int64 or a double).
float ceil_to_float(long double x) {
  float f = (float)x;
  if (!isnan(f) && (long double)f < x) f = nextafterf(f, INFINITY);
  return f;
}
double ceil_to_double(long double x) {
  double d = (double)x;
  if (!isnan(d) && (long double)d < x) d = nextafter(d, INFINITY);
  return d;
}

}  // namespace

namespace msh {

// Resolve Julia's promotion for this call (src/marching_cubes.jl:31,
// :100-104; src/marching_tetrahedra.jl:43-54,131) into a Plan.
int make_plan(std::string* c, const msh_field* f, const msh_params* p, Plan* out) {
  if (!f || !p) return fail(c, MSH_ERR_ARGUMENT, "null field or params");
  if (f->dtype != K_F32 && f->dtype != K_F64)
    return fail(c, MSH_ERR_UNSUPPORTED, "field element type must be Float32 or Float64");
  if (f->memory != MSH_HOST && f->memory != MSH_DEVICE)
    return fail(c, MSH_ERR_ARGUMENT, "field memory must be MSH_HOST or MSH_DEVICE");
  if (f->nx < 0 || f->ny < 0 || f->nz < 0)
    return fail(c, MSH_ERR_ARGUMENT, "negative array size");
  if (f->nx > INT32_MAX / 2 || f->ny > INT32_MAX / 2 || f->nz > INT32_MAX / 2)
    return fail(c, MSH_ERR_ARGUMENT, "array dimension too large");
  // the kernels address one z-plane with 32-bit byte offsets
  if ((double)f->nx * (double)f->ny * (f->dtype == K_F64 ? 8.0 : 4.0) >= 2147483648.0)
    return fail(c, MSH_ERR_UNSUPPORTED, "one z-plane of the field must be smaller than 2 GiB");
  if (p->method != MSH_MARCHING_CUBES && p->method != MSH_MARCHING_TETRAHEDRA &&
      p->method != MSH_NAIVE_SURFACE_NETS)
    return fail(c, MSH_ERR_ARGUMENT, "unknown meshing algorithm");
  if (!valid_kind(p->iso_kind) || !valid_kind(p->axis_kind[0]) || !valid_kind(p->axis_kind[1]) ||
      !valid_kind(p->axis_kind[2]) ||
      (p->method == MSH_MARCHING_TETRAHEDRA && !valid_kind(p->eps_kind)))
    return fail(c, MSH_ERR_UNSUPPORTED, "iso, eps and axis types must be Int64, Float32 or Float64");
  if (p->iso_kind == K_INT64 && (p->iso_int > (1ll << 53) || p->iso_int < -(1ll << 53)))
    return fail(c, MSH_ERR_UNSUPPORTED, "integer iso beyond 2^53 is not supported");

  Plan pl{};
  pl.method = p->method;
  pl.T = f->dtype;
  const int I = p->iso_kind;
  pl.M = (pl.T == K_F64 || I == K_F64) ? K_F64 : K_F32;
  const bool mt = p->method == MSH_MARCHING_TETRAHEDRA;
  const int E = mt ? p->eps_kind : K_INT64;
  pl.A = (pl.M == K_F64 || E == K_F64) ? K_F64 : K_F32;
  const bool any64 = pl.T == K_F64 || I == K_F64 || E == K_F64 || p->axis_kind[0] == K_F64 ||
                     p->axis_kind[1] == K_F64 || p->axis_kind[2] == K_F64;
  pl.Out = any64 ? K_F64 : K_F32;

  // iso in M, and the classification threshold in T
  long double iso_exact;
  if (I == K_INT64) {
    iso_exact = (long double)p->iso_int;
    pl.isoM = pl.M == K_F32 ? (double)(float)p->iso_int : (double)p->iso_int;
  } else if (I == K_F32) {
    iso_exact = (long double)(float)p->iso;
    pl.isoM = (double)(float)p->iso;
  } else {
    iso_exact = (long double)p->iso;
    pl.isoM = p->iso;
  }
  if (I != K_INT64 && isnan(p->iso)) {
    pl.thr = NAN;
  } else {
    pl.thr = pl.T == K_F32 ? (double)ceil_to_float(iso_exact) : ceil_to_double(iso_exact);
  }

  // MT: eps in A and one(T) - eps in promote(T, E)
  if (mt) {
    if (E == K_INT64) {
      pl.epsA = pl.A == K_F32 ? (double)(float)p->eps_int : (double)p->eps_int;
      pl.omeA = pl.T == K_F32 ? (double)(1.0f - (float)p->eps_int) : 1.0 - (double)p->eps_int;
    } else if (E == K_F32) {
      const float e = (float)p->eps;
      pl.epsA = (double)e;
      pl.omeA = pl.T == K_F32 ? (double)(1.0f - e) : 1.0 - (double)e;
    } else {
      pl.epsA = p->eps;
      pl.omeA = 1.0 - p->eps;
    }
  }

  // LinRange(first(X), last(X), n): element type Float32 for Float32 ends,
  // Float64 otherwise; length 1 requires equal ends (Julia ArgumentError).
  const int64_t n[3] = {f->nx, f->ny, f->nz};
  for (int k = 0; k < 3; ++k) {
    pl.cx[k] = p->axis_kind[k] == K_F32 ? K_F32 : K_F64;
    if (pl.cx[k] == K_F32) {
      pl.a[k] = (double)(float)p->first[k];
      pl.b[k] = (double)(float)p->last[k];
    } else {
      pl.a[k] = p->first[k];
      pl.b[k] = p->last[k];
    }
    if (n[k] == 1 && !(pl.a[k] == pl.b[k]))
      return fail(c, MSH_ERR_ARGUMENT,
                  "ArgumentError: range(" + std::to_string(pl.a[k]) + ", stop=" +
                      std::to_string(pl.b[k]) + ", length=1): endpoints differ");
  }
  *out = pl;
  return MSH_OK;
}

// whole: z rows are split only to keep units within 4096 cells (a batch,
// whose launch has blocks enough from its many volumes, and the entry point
// without allocation, whose fixed workspace then holds the most rows)
Geom make_geom(const msh_field* f, int method, bool whole) {
  Geom g{};
  g.nx = (int32_t)f->nx;
  g.ny = (int32_t)f->ny;
  g.nz = (int32_t)f->nz;
  const int64_t cx = f->nx - 1, cy = f->ny - 1, cz = f->nz - 1;
  if (cx <= 0 || cy <= 0 || cz <= 0) {
    g.zseg = 1;
    g.nzseg = 0;
    g.units = 0;
    return g;
  }
  // Split z rows into segments so that the count has about `target` tiles
  // (252 x-cells by 6-8 rows each) and the emit about as many more as it
  // has per tile layer: an emit tile walks one segment, and more, shorter
  // tiles even out the last round of blocks on the CUs at the price of a
  // set-up per tile.  Segments are whole multiples of `gran` cells; at most
  // 4096 cells keep the packed 16-bit per-unit counts from overflowing.
  // Measured (Float32): Marching Cubes at 1024 tiles and 64-cell steps
  // 1024^3 5.77 -> 5.64 ms and Float64 8.96 -> 8.85 ms (two segments), 512^3
  // 1.29 -> 1.17 ms, 256^3 unchanged (with 32-cell steps it lost, 0.100 ->
  // 0.111 ms); tetrahedra and surface nets at 2048 tiles and 32-cell steps 256^3
  // 0.220 -> 0.213 and 0.165 -> 0.155 ms, 512^3 2.59 -> 2.32 and 1.46 ->
  // 1.29 ms, 128^3 tetrahedra 0.276 -> 0.240 ms (against 512 tiles and
  // 64-cell steps for every method).
  const int64_t ty = f->dtype == K_F64 ? 6 : 8;
  const int64_t blocks_xy = ((cx + 251) / 252) * ((cy + ty - 1) / ty);
  const int64_t target = method == MC ? 1024 : 2048;
  const int64_t gran = method == MC ? 64 : 32;
  int64_t nseg = whole ? 1 : (target + blocks_xy - 1) / blocks_xy;
  const int64_t max_seg = (cz + gran - 1) / gran;
  if (nseg > max_seg) nseg = max_seg;
  if (nseg < 1) nseg = 1;
  const int64_t min_seg = (cz + 4095) / 4096;
  if (nseg < min_seg) nseg = min_seg;
  int64_t zseg = (cz + nseg - 1) / nseg;
  zseg = (zseg + gran - 1) / gran * gran;
  if (zseg > 4096) zseg = 4096;
  nseg = (cz + zseg - 1) / zseg;
  g.zseg = (int32_t)zseg;
  g.nzseg = (int32_t)nseg;
  g.units = cx * cy * nseg;
  return g;
}

// Workspace of `units` units and `ncoords` coordinates placed at `base`
// (null: only the size); returns the bytes it takes.
size_t workspace_layout(int64_t units, int64_t ncoords, void* base, Workspace* w) {
  const int64_t U = units > 0 ? units : 1;
  const int64_t nt = scan_tiles(U);
  const size_t b_cnt = align256((size_t)U * 4);
  const size_t b_off = align256((size_t)U * 8);  // also the z masks
  const size_t b_tsum = align256((size_t)nt * 8);
  const size_t b_toff = align256((size_t)nt * 16);
  const size_t b_tot = 256;
  const size_t b_coords = align256((size_t)(ncoords + 1) * 8);
  const size_t need = b_cnt + 3 * b_off + b_tsum + b_toff + b_tot + b_coords;
  if (!base || !w) return need;
  char* p = (char*)base;
  w->cnt = (uint32_t*)p;
  p += b_cnt;
  w->zmask = (uint64_t*)p;
  p += b_off;
  w->voff = (int64_t*)p;
  p += b_off;
  w->foff = (int64_t*)p;
  p += b_off;
  w->tilesum = (uint64_t*)p;
  p += b_tsum;
  w->tileoff = (int64_t*)p;
  p += b_toff;
  w->totals = (int64_t*)p;
  p += b_tot;
  w->coords = (double*)p;
  return need;
}

hipError_t tables_setup(int device, hipStream_t st) {
  static std::mutex mu;
  static bool done[64];
  if (device < 0 || device >= 64) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (done[device]) return hipSuccess;
  // the caller (msh_create) has made `device` current
  hipError_t e = launch_count_tables(st);
  if (e == hipSuccess) e = launch_mt_tables(st);
  // complete before any context's stream can launch a kernel that reads them
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) done[device] = true;
  return e;
}

}  // namespace msh

namespace {

int make_plan(Ctx* c, const msh_field* f, const msh_params* p, Plan* out) {
  return msh::make_plan(c ? &c->err : nullptr, f, p, out);
}

int grow(Ctx* c, void** buf, size_t* have, size_t need, const char* what) {
  if (need <= *have) return MSH_OK;
  if (*buf) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
  }
  hipError_t e = hipMalloc(buf, need);
  if (e != hipSuccess) {
    *buf = nullptr;
    return hip_fail(c, e, what);
  }
  *have = need;
  return MSH_OK;
}

int setup_workspace(Ctx* c, int64_t units, int64_t ncoords, Workspace* w) {
  const size_t need = workspace_layout(units, ncoords, nullptr, nullptr);
  int rc = grow(c, &c->ws, &c->ws_bytes, need, "workspace allocation");
  if (rc) return rc;
  workspace_layout(units, ncoords, c->ws, w);
  return MSH_OK;
}

bool same_call(const Ctx* c, const msh_field* f, const msh_params* p) {
  return c->have && memcmp(&c->field, f, sizeof(*f)) == 0 &&
         memcmp(&c->params, p, sizeof(*p)) == 0;
}

bool same_batch(const Ctx* c, int32_t n, const msh_field* f, const msh_params* p) {
  const BatchState& b = c->batch;
  return b.have && (size_t)n == b.fields.size() &&
         (n == 0 || (memcmp(b.fields.data(), f, sizeof(*f) * (size_t)n) == 0 &&
                     memcmp(b.params.data(), p, sizeof(*p) * (size_t)n) == 0));
}

// The host copy of the descriptors may be rewritten only once the previous
// upload from it has run.
int wait_upload(Ctx* c) {
  if (!c->upload_pending) return MSH_OK;
  hipError_t e = hipEventSynchronize(c->upload);
  if (e != hipSuccess) return hip_fail(c, e, "descriptor upload");
  c->upload_pending = false;
  return MSH_OK;
}

bool same_plan_types(const Plan& a, const Plan& b) {
  return a.method == b.method && a.T == b.T && a.M == b.M && a.A == b.A && a.Out == b.Out;
}

}  // namespace

extern "C" {

int msh_create(int device, void** out) {
  if (!out) return MSH_ERR_ARGUMENT;
  *out = nullptr;
  Ctx* c = new (std::nothrow) Ctx();
  if (!c) return MSH_ERR_OUT_OF_MEMORY;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&c->pinned, 4 * sizeof(int64_t), hipHostMallocDefault);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&c->upload, hipEventDisableTiming);
  // once per device, with the device current: the address of its copy of
  // the static scratch array
  if (e == hipSuccess) e = noalloc_setup(device, &c->scratch);
  if (e == hipSuccess) e = tables_setup(device, c->own);
  if (e != hipSuccess) {
    if (c->upload) (void)hipEventDestroy(c->upload);
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
    return e == hipErrorOutOfMemory ? MSH_ERR_OUT_OF_MEMORY : MSH_ERR_DEVICE;
  }
  c->device = device;
  c->stream = c->own;
  *out = c;
  return MSH_OK;
}

int msh_destroy(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->ws) (void)hipFree(c->ws);
  if (c->fstage) (void)hipFree(c->fstage);
  if (c->ostage) (void)hipFree(c->ostage);
  if (c->bws) (void)hipFree(c->bws);
  if (c->wws) (void)hipFree(c->wws);
  if (c->hbatch) (void)hipHostFree(c->hbatch);
  graph_drop(&c->gcount);
  graph_drop(&c->gemit);
  for (int i = 0; i < kSpreadWays - 1; ++i) {
    if (c->aux[i]) (void)hipStreamSynchronize(c->aux[i]);
    if (c->aux[i]) (void)hipStreamDestroy(c->aux[i]);
    if (c->join[i]) (void)hipEventDestroy(c->join[i]);
  }
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->upload) (void)hipEventDestroy(c->upload);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->own) (void)hipStreamDestroy(c->own);
  delete c;
  return MSH_OK;
}

int msh_set_stream(void* ctx, void* s) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  // NULL is HIP's default (null) stream, which is what torch reports for
  // its default stream; it is used as given, never replaced by our own.
  c->stream = (hipStream_t)s;
  return MSH_OK;
}

int msh_synchronize(void* ctx) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  hipError_t e = hipStreamSynchronize(c->stream);
  return e == hipSuccess ? MSH_OK : hip_fail(c, e, "hipStreamSynchronize");
}

const char* msh_last_error(const void* ctx) {
  const Ctx* c = (const Ctx*)ctx;
  return c ? c->err.c_str() : "null context";
}

int msh_output_kind(const msh_field* f, const msh_params* p, int32_t* kind) {
  if (!f || !p || !kind) return MSH_ERR_ARGUMENT;
  Plan pl;
  msh_field f1 = *f;
  // the kind does not depend on the sizes; avoid the length-1 range check
  f1.nx = f1.ny = f1.nz = 2;
  int rc = msh::make_plan(nullptr, &f1, p, &pl);
  if (rc) return rc;
  *kind = pl.Out;
  return MSH_OK;
}

int msh_count(void* ctx, const msh_field* f, const msh_params* p, int64_t* nverts,
              int64_t* nfaces) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  c->have = false;
  c->batch.have = false;  // the batch state shares the workspace
  Plan pl;
  int rc = make_plan(c, f, p, &pl);
  if (rc) return rc;
  pl.family = c->family[pl.method][pl.T][size_class(f)];
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  Geom g = make_geom(f, pl.method);
  Workspace w{};
  rc = setup_workspace(c, g.units, (int64_t)g.nx + g.ny + g.nz, &w);
  if (rc) return rc;
  int64_t nv = 0, nfc = 0;
  const void* dfield = f->data;
  if (g.units > 0) {
    const size_t esz = f->dtype == K_F32 ? 4 : 8;
    const size_t fbytes = (size_t)f->nx * f->ny * f->nz * esz;
    if (!f->data) return fail(c, MSH_ERR_ARGUMENT, "null field data");
    if (f->memory == MSH_HOST) {
      rc = grow(c, &c->fstage, &c->fstage_bytes, fbytes, "field staging allocation");
      if (rc) return rc;
      hipError_t e = hipMemcpyAsync(c->fstage, f->data, fbytes, hipMemcpyHostToDevice, c->stream);
      if (e != hipSuccess) return hip_fail(c, e, "field upload");
      dfield = c->fstage;
    }
    const VolArgs a = make_vol(pl, g, dfield, w);
    hipError_t e = launch_count(pl, a, c->stream);
    // the scan also fills the coordinate tables and writes the sizes into
    // the pinned host words itself; a negative word afterwards (not
    // written) is read back by a copy
    c->pinned[1] = c->pinned[3] = -1;
    const CoordArgs ca = coord_args(pl, g, w.coords);
    if (e == hipSuccess) e = launch_scan(g, w, c->stream, c->pinned, &ca);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && (c->pinned[1] < 0 || c->pinned[3] < 0)) {
      e = hipMemcpyAsync(c->pinned, w.totals, 4 * sizeof(int64_t), hipMemcpyDeviceToHost,
                         c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    }
    if (e != hipSuccess) return hip_fail(c, e, "count/scan");
    nv = c->pinned[1];
    nfc = c->pinned[3];
    c->vol = a;
  }
  c->field = *f;
  c->params = *p;
  c->plan = pl;
  c->geom = g;
  c->w = w;
  c->nverts = nv;
  c->nfaces = nfc;
  c->have = true;
  if (nverts) *nverts = nv;
  if (nfaces) *nfaces = nfc;
  return MSH_OK;
}

int msh_extract(void* ctx, const msh_field* f, const msh_params* p, void* verts,
                int64_t* faces, int32_t out_memory) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (!f || !p) return fail(c, MSH_ERR_ARGUMENT, "null field or params");
  if (!same_call(c, f, p))
    return fail(c, MSH_ERR_STATE, "msh_extract must follow msh_count with the same arguments");
  if (out_memory != MSH_HOST && out_memory != MSH_DEVICE)
    return fail(c, MSH_ERR_ARGUMENT, "out_memory must be MSH_HOST or MSH_DEVICE");
  if (c->nverts == 0 && c->nfaces == 0) return MSH_OK;
  if ((c->nverts > 0 && !verts) || (c->nfaces > 0 && !faces))
    return fail(c, MSH_ERR_ARGUMENT, "null output buffer");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  const size_t osz = c->plan.Out == K_F32 ? 4 : 8;
  const size_t vbytes = (size_t)c->nverts * 3 * osz, fbytes = (size_t)c->nfaces * 3 * 8;
  void* dv = verts;
  int64_t* df = faces;
  if (out_memory == MSH_HOST) {
    int rc = grow(c, &c->ostage, &c->ostage_bytes, align256(vbytes) + fbytes + 256,
                  "output staging allocation");
    if (rc) return rc;
    dv = c->ostage;
    df = (int64_t*)((char*)c->ostage + align256(vbytes));
  }
  VolArgs a = c->vol;
  a.verts = dv;
  a.faces = df;
  hipError_t e = launch_emit(c->plan, a, c->stream);
  if (e == hipErrorNotSupported)
    return fail(c, MSH_ERR_UNSUPPORTED, "this method is not available on the GPU path yet");
  if (e != hipSuccess) return hip_fail(c, e, "emit");
  if (out_memory == MSH_HOST) {
    if (vbytes) e = hipMemcpyAsync(verts, dv, vbytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && fbytes)
      e = hipMemcpyAsync(faces, df, fbytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "output download");
  }
  return MSH_OK;
}

int msh_isosurface_noalloc(void* ctx, const msh_field* f, const msh_params* p, void* verts,
                           int64_t vcap, int64_t* faces, int64_t fcap, int64_t* sizes) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  Plan pl;
  int rc = make_plan(c, f, p, &pl);
  if (rc) return rc;
  pl.family = c->family[pl.method][pl.T][size_class(f)];
  if (f->memory != MSH_DEVICE)
    return fail(c, MSH_ERR_ARGUMENT, "msh_isosurface_noalloc takes a device field");
  if (vcap < 0 || fcap < 0) return fail(c, MSH_ERR_ARGUMENT, "negative output capacity");
  if ((vcap > 0 && !verts) || (fcap > 0 && !faces))
    return fail(c, MSH_ERR_ARGUMENT, "null output buffer");
  if (!c->scratch) return fail(c, MSH_ERR_DEVICE, "no static scratch on this device");
  const Geom g = make_geom(f, pl.method, true);
  // the workspace, then the clamped sizes the emit kernels read
  const size_t need = workspace_layout(g.units, (int64_t)g.nx + g.ny + g.nz, nullptr, nullptr);
  if (need + 256 > kScratchBytes)
    return fail(c, MSH_ERR_ARGUMENT,
                "volume needs " + std::to_string(need + 256) + " bytes of workspace, more than "
                "the " + std::to_string(kScratchBytes) + "-byte static scratch");
  if (g.units > 0 && !f->data) return fail(c, MSH_ERR_ARGUMENT, "null field data");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  Workspace w{};
  workspace_layout(g.units, (int64_t)g.nx + g.ny + g.nz, c->scratch, &w);
  int64_t* clamp = (int64_t*)((char*)c->scratch + need);
  if (g.units == 0) {  // no cells: nothing to write, zero sizes
    const hipError_t e = launch_noalloc_sizes(nullptr, vcap, fcap, clamp, sizes, c->stream);
    return e == hipSuccess ? MSH_OK : hip_fail(c, e, "sizes");
  }
  VolArgs a = make_vol(pl, g, f->data, w);
  hipError_t e = launch_coords(pl, g, w.coords, c->stream);
  if (e == hipSuccess) e = launch_count(pl, a, c->stream);
  if (e == hipSuccess) e = launch_scan(g, w, c->stream);
  if (e == hipSuccess) e = launch_noalloc_sizes(w.totals, vcap, fcap, clamp, sizes, c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "count/scan");
  a.sz = clamp;
  a.verts = verts;
  a.faces = faces;
  e = launch_emit(pl, a, c->stream);
  if (e == hipErrorNotSupported)
    return fail(c, MSH_ERR_UNSUPPORTED, "this method is not available on the GPU path yet");
  if (e != hipSuccess) return hip_fail(c, e, "emit");
  return MSH_OK;
}

int msh_batch_count(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                    int64_t* nverts, int64_t* nfaces) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  BatchState& b = c->batch;
  b.have = false;
  c->have = false;  // the single-volume state shares the workspace
  if (n < 0) return fail(c, MSH_ERR_ARGUMENT, "negative batch size");
  if (n > 0 && (!fields || !params)) return fail(c, MSH_ERR_ARGUMENT, "null fields or params");
  int rc = wait_upload(c);
  if (rc) return rc;
  b.plans.assign((size_t)n, Plan{});
  std::vector<Geom> geoms((size_t)n);
  for (int v = 0; v < n; ++v) {
    rc = make_plan(c, &fields[v], &params[v], &b.plans[v]);
    if (rc) {
      c->err = "volume " + std::to_string(v) + ": " + c->err;
      return rc;
    }
    b.plans[v].family = c->family[b.plans[v].method][b.plans[v].T][size_class(&fields[v])];
    geoms[v] = make_geom(&fields[v], b.plans[v].method, true);
    if (geoms[v].units > 0 && !fields[v].data)
      return fail(c, MSH_ERR_ARGUMENT, "volume " + std::to_string(v) + ": null field data");
  }
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");

  // units, coordinates and host staging, in volume order
  int64_t units = 0, ncoords = 0;
  size_t hbytes = 0;
  std::vector<size_t> hoff((size_t)n, 0);
  for (int v = 0; v < n; ++v) {
    units += geoms[v].units;
    ncoords += (int64_t)geoms[v].nx + geoms[v].ny + geoms[v].nz;
    if (fields[v].memory == MSH_HOST && geoms[v].units > 0) {
      hoff[v] = hbytes;
      hbytes += align256((size_t)fields[v].nx * fields[v].ny * fields[v].nz *
                         (fields[v].dtype == K_F32 ? 4 : 8));
    }
  }
  Workspace w{};
  rc = setup_workspace(c, units, ncoords, &w);
  if (rc) return rc;
  if (hbytes) {
    rc = grow(c, &c->fstage, &c->fstage_bytes, hbytes, "field staging allocation");
    if (rc) return rc;
  }

  // descriptors; count launches group volumes by (T, method), emit launches
  // by the full instantiation (types + variant)
  b.vols.assign((size_t)n, VolArgs{});
  std::vector<int> emit_var((size_t)n, 0);
  int64_t ubase = 0, cbase = 0;
  for (int v = 0; v < n; ++v) {
    const void* df = fields[v].data;
    if (fields[v].memory == MSH_HOST && geoms[v].units > 0) df = (char*)c->fstage + hoff[v];
    VolArgs a = make_vol(b.plans[v], geoms[v], df, w);
    count_tiling(b.plans[v], &a, false);
    a.coords = w.coords + cbase;
    a.ubase = ubase;
    a.sz = nullptr;  // set below, once the sizes array is placed
    emit_var[v] = emit_variant(b.plans[v], &a);
    if (geoms[v].units == 0) {
      a.cntiles = 0;
      a.ntiles = 0;
    }
    b.vols[v] = a;
    ubase += geoms[v].units;
    cbase += (int64_t)geoms[v].nx + geoms[v].ny + geoms[v].nz;
  }
  b.counts.clear();
  b.emits.clear();
  // Within a launch, the volumes' tiles are queued longest first (a tile
  // walks the volume's z extent), so the launch does not end waiting on a
  // long tile taken last.
  std::vector<int> ord((size_t)n);
  for (int v = 0; v < n; ++v) ord[v] = v;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](int x, int y) { return geoms[x].nz > geoms[y].nz; });
  std::vector<char> done((size_t)n, 0);
  int64_t cblocks = 0;
  for (int v : ord) {
    if (done[v]) continue;
    Launch L;
    L.plan = b.plans[v];
    L.gbase = cblocks;
    for (int u : ord) {
      if (done[u] || b.plans[u].T != L.plan.T || b.plans[u].method != L.plan.method) continue;
      done[u] = 1;
      b.vols[u].cblk0 = cblocks;
      cblocks += b.vols[u].cntiles;
      L.cost += cell_cost(b.vols[u], b.plans[u], false);
    }
    L.nblocks = cblocks - L.gbase;
    if (L.nblocks) b.counts.push_back(L);
  }
  done.assign((size_t)n, 0);
  int64_t eblocks = 0;
  for (int v : ord) {
    if (done[v]) continue;
    Launch L;
    L.plan = b.plans[v];
    L.variant = emit_var[v];
    L.gbase = eblocks;
    for (int u : ord) {
      if (done[u] || emit_var[u] != L.variant || !same_plan_types(b.plans[u], L.plan)) continue;
      done[u] = 1;
      b.vols[u].eblk0 = eblocks;
      eblocks += b.vols[u].ntiles;
      L.cost += cell_cost(b.vols[u], b.plans[u], true);
    }
    L.nblocks = eblocks - L.gbase;
    if (L.nblocks) b.emits.push_back(L);
  }
  if (cblocks > INT32_MAX || eblocks > INT32_MAX)
    return fail(c, MSH_ERR_ARGUMENT, "batch too large for one launch");

  // device side: descriptors | cmap | emap | sizes
  const size_t b_vols = align256(sizeof(VolArgs) * (size_t)(n > 0 ? n : 1));
  const size_t b_cmap = align256(4 * (size_t)(cblocks + 1));
  const size_t b_emap = align256(4 * (size_t)(eblocks + 1));
  const size_t b_sizes = align256(32 * (size_t)(n > 0 ? n : 1));
  const size_t nctr = 8 * (b.counts.size() + b.emits.size());
  const size_t b_ctr = align256(4 * (nctr + 1));
  rc = grow(c, &c->bws, &c->bws_bytes, b_vols + b_cmap + b_emap + b_sizes + b_ctr,
            "batch workspace");
  if (rc) return rc;
  VolArgs* dvols = (VolArgs*)c->bws;
  int32_t* cmap = (int32_t*)((char*)c->bws + b_vols);
  int32_t* emap = (int32_t*)((char*)cmap + b_cmap);
  int64_t* dsizes = (int64_t*)((char*)emap + b_emap);
  b.ctr_off = b_vols + b_cmap + b_emap + b_sizes;
  uint32_t* ctrs = (uint32_t*)((char*)c->bws + b.ctr_off);
  for (int v = 0; v < n; ++v) b.vols[v].sz = dsizes + 4 * v;
  // pinned host side: descriptors | sizes
  if (c->hbatch_bytes < b_vols + b_sizes) {
    if (c->hbatch) (void)hipHostFree(c->hbatch);
    c->hbatch = nullptr;
    c->hbatch_bytes = 0;
    hipError_t e = hipHostMalloc(&c->hbatch, b_vols + b_sizes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(c, e, "pinned batch buffer");
    c->hbatch_bytes = b_vols + b_sizes;
  }
  int64_t* hsizes = (int64_t*)((char*)c->hbatch + b_vols);
  if (n > 0) memcpy(c->hbatch, b.vols.data(), sizeof(VolArgs) * (size_t)n);

  bool host_fields = false;
  for (int v = 0; v < n; ++v) host_fields |= fields[v].memory == MSH_HOST && geoms[v].units > 0;
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = hipSuccess;
    for (int v = 0; v < n && e == hipSuccess; ++v)
      if (fields[v].memory == MSH_HOST && geoms[v].units > 0)
        e = hipMemcpyAsync((char*)c->fstage + hoff[v], fields[v].data,
                           (size_t)fields[v].nx * fields[v].ny * fields[v].nz *
                               (fields[v].dtype == K_F32 ? 4 : 8),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && n > 0)
      e = hipMemcpyAsync(dvols, c->hbatch, sizeof(VolArgs) * (size_t)n, hipMemcpyHostToDevice,
                         c->stream);
    if (e == hipSuccess) e = launch_fill_maps(dvols, n, cmap, emap, c->stream);
    if (e == hipSuccess) e = launch_coords_batch(dvols, n, c->stream);
    if (e == hipSuccess) e = launch_zero_u32(ctrs, (int64_t)(8 * b.counts.size()), c->stream);
    Spread sp{c};
    if (e == hipSuccess) e = sp.start((int)b.counts.size());
    for (size_t k : by_cost(b.counts)) {
      if (e != hipSuccess) break;
      e = launch_count_batch(b.counts[k].plan, dvols, cmap, b.counts[k].gbase,
                             b.counts[k].nblocks, ctrs + 8 * k, sp.next(b.counts[k].cost));
    }
    const hipError_t ej = sp.finish();
    if (e == hipSuccess) e = ej;
    Geom gall{};
    gall.units = units;
    if (e == hipSuccess && units > 0) e = launch_scan(gall, w, c->stream);
    if (e == hipSuccess && units > 0)
      e = launch_sizes_batch(dvols, n, w, units, dsizes, c->stream);
    if (e == hipSuccess && units > 0)
      e = hipMemcpyAsync(hsizes, dsizes, 32 * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    return e;
  };
  // everything the launches and copies above are derived from: the
  // arguments, the kernel family choices and the buffers
  std::vector<char> key;
  key_add(&key, fields, (size_t)n);
  key_add(&key, params, (size_t)n);
  key_add(&key, &c->family[0][0][0], 9 * kSizeClasses);
  const void* bufs[5] = {c->ws, c->bws, c->hbatch, c->fstage, (const void*)c->stream};
  key_add(&key, bufs, 5);
  hipError_t e = run_phase(c, &c->gcount, std::move(key), !host_fields && c->stream != nullptr,
                           enqueue);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "batched count/scan");

  b.sizes.assign(4 * (size_t)n, 0);
  if (units > 0) memcpy(b.sizes.data(), hsizes, 32 * (size_t)n);
  for (int v = 0; v < n; ++v) {
    if (nverts) nverts[v] = b.sizes[4 * v + 1];
    if (nfaces) nfaces[v] = b.sizes[4 * v + 3];
  }
  b.fields.assign(fields, fields + n);
  b.params.assign(params, params + n);
  b.units = units;
  b.ncoords = ncoords;
  b.cblocks = cblocks;
  b.eblocks = eblocks;
  b.have = true;
  return MSH_OK;
}

int msh_batch_extract(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                      void* const* verts, int64_t* const* faces, int32_t out_memory) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  BatchState& b = c->batch;
  if (n < 0 || (n > 0 && (!fields || !params)))
    return fail(c, MSH_ERR_ARGUMENT, "null fields or params");
  if (!same_batch(c, n, fields, params))
    return fail(c, MSH_ERR_STATE,
                "msh_batch_extract must follow msh_batch_count with the same arguments");
  if (out_memory != MSH_HOST && out_memory != MSH_DEVICE)
    return fail(c, MSH_ERR_ARGUMENT, "out_memory must be MSH_HOST or MSH_DEVICE");
  if (n == 0 || b.eblocks == 0) return MSH_OK;
  if (!verts || !faces) return fail(c, MSH_ERR_ARGUMENT, "null output pointer arrays");
  for (int v = 0; v < n; ++v)
    if ((b.sizes[4 * v + 1] > 0 && !verts[v]) || (b.sizes[4 * v + 3] > 0 && !faces[v]))
      return fail(c, MSH_ERR_ARGUMENT, "volume " + std::to_string(v) + ": null output buffer");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  int rc = wait_upload(c);
  if (rc) return rc;

  std::vector<size_t> vb((size_t)n), fb((size_t)n), off((size_t)n);
  size_t stage = 0;
  for (int v = 0; v < n; ++v) {
    const size_t osz = b.plans[v].Out == K_F32 ? 4 : 8;
    vb[v] = (size_t)b.sizes[4 * v + 1] * 3 * osz;
    fb[v] = (size_t)b.sizes[4 * v + 3] * 3 * 8;
    off[v] = stage;
    stage += align256(vb[v]) + align256(fb[v]);
  }
  if (out_memory == MSH_HOST) {
    rc = grow(c, &c->ostage, &c->ostage_bytes, stage + 256, "output staging allocation");
    if (rc) return rc;
  }
  for (int v = 0; v < n; ++v) {
    if (out_memory == MSH_HOST) {
      b.vols[v].verts = (char*)c->ostage + off[v];
      b.vols[v].faces = (int64_t*)((char*)c->ostage + off[v] + align256(vb[v]));
    } else {
      b.vols[v].verts = verts[v];
      b.vols[v].faces = faces[v];
    }
  }
  VolArgs* dvols = (VolArgs*)c->bws;
  const size_t b_vols = align256(sizeof(VolArgs) * (size_t)n);
  const size_t b_cmap = align256(4 * (size_t)(b.cblocks + 1));
  int32_t* emap = (int32_t*)((char*)c->bws + b_vols + b_cmap);
  memcpy(c->hbatch, b.vols.data(), sizeof(VolArgs) * (size_t)n);
  uint32_t* ectr = (uint32_t*)((char*)c->bws + b.ctr_off) + 8 * b.counts.size();
  // the output pointers travel in the descriptors (copied from hbatch by
  // the first operation), so a captured emit phase serves new outputs
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(dvols, c->hbatch, sizeof(VolArgs) * (size_t)n,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipEventRecord(c->upload, c->stream);
    if (e == hipSuccess) e = launch_zero_u32(ectr, (int64_t)(8 * b.emits.size()), c->stream);
    Spread sp{c};
    if (e == hipSuccess) e = sp.start((int)b.emits.size());
    for (size_t k : by_cost(b.emits)) {
      if (e != hipSuccess) break;
      e = launch_emit_batch(b.emits[k].plan, b.emits[k].variant, dvols, emap, b.emits[k].gbase,
                            b.emits[k].nblocks, ectr + 8 * k, sp.next(b.emits[k].cost));
    }
    const hipError_t ej = sp.finish();
    if (e == hipSuccess) e = ej;
    return e;
  };
  std::vector<char> key;
  key_add(&key, b.fields.data(), (size_t)n);
  key_add(&key, b.params.data(), (size_t)n);
  key_add(&key, &c->family[0][0][0], 9 * kSizeClasses);
  const void* bufs[5] = {c->ws, c->bws, c->hbatch, (const void*)c->stream,
                         (const void*)c->upload};
  key_add(&key, bufs, 5);
  hipError_t e = run_phase(c, &c->gemit, std::move(key),
                           out_memory == MSH_DEVICE && c->stream != nullptr, enqueue);
  if (e == hipSuccess) c->upload_pending = true;
  if (e != hipSuccess) return hip_fail(c, e, "batched emit");
  if (out_memory == MSH_HOST) {
    for (int v = 0; v < n && e == hipSuccess; ++v) {
      if (vb[v]) e = hipMemcpyAsync(verts[v], b.vols[v].verts, vb[v], hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess && fb[v])
        e = hipMemcpyAsync(faces[v], b.vols[v].faces, fb[v], hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, e, "output download");
  }
  return MSH_OK;
}

int msh_batch_graph_launches(void* ctx, int64_t* count_phase, int64_t* emit_phase) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (count_phase) *count_phase = c->gcount.replays;
  if (emit_phase) *emit_phase = c->gemit.replays;
  return MSH_OK;
}

int msh_set_kernel_family(void* ctx, int32_t method, int32_t dtype, int32_t family) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (method < 0 || method > 2 || (dtype != K_F32 && dtype != K_F64) || family < -1 ||
      family > 3)
    return fail(c, MSH_ERR_ARGUMENT, "method, dtype or kernel family out of range");
  for (int& f : c->family[method][dtype]) f = family;
  c->have = false;
  c->batch.have = false;
  return MSH_OK;
}

int msh_autotune(void* ctx, const msh_field* f, const msh_params* p, int32_t* family) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  int64_t nv = 0, nf = 0;
  int rc = msh_count(ctx, f, p, &nv, &nf);
  if (rc) return rc;
  const Plan base = c->plan;
  const int cls = size_class(f);
  if (family) *family = c->family[base.method][base.T][cls];
  if (nv == 0 && nf == 0) return MSH_OK;  // nothing to time
  const size_t osz = base.Out == K_F32 ? 4 : 8;
  const size_t vbytes = align256((size_t)nv * 3 * osz), fbytes = (size_t)nf * 3 * 8;
  rc = grow(c, &c->ostage, &c->ostage_bytes, vbytes + fbytes + 256, "autotune output buffer");
  if (rc) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int best = -1;
  float best_ms = 0.f;
  for (int fam = 3; fam >= 0 && e == hipSuccess; --fam) {
    Plan q = base;
    q.family = fam;
    VolArgs a = c->vol;
    a.verts = c->ostage;
    a.faces = (int64_t*)((char*)c->ostage + vbytes);
    VolArgs probe = a;
    if (variant_family(emit_variant(q, &probe)) != fam) continue;  // not for this volume
    e = launch_emit(q, a, c->stream);  // warm-up
    // the best of at least 3 launches, and of as many as fill about 2 ms
    // (short launches vary more)
    float ms = 0.f;
    int reps = 3;
    for (int rep = 0; rep < reps && e == hipSuccess; ++rep) {
      e = hipEventRecord(e0, c->stream);
      if (e == hipSuccess) e = launch_emit(q, a, c->stream);
      if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e == hipSuccess && (rep == 0 || t < ms)) ms = t;
      if (rep == 0 && t > 0.f) reps = t >= 2.f / 3 ? 3 : (int)(2.f / t) < 50 ? (int)(2.f / t) : 50;
    }
    if (e == hipSuccess && (best < 0 || ms < best_ms)) {
      best = fam;
      best_ms = ms;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(c, e, "autotune");
  c->family[base.method][base.T][cls] = best;
  c->plan.family = best;
  if (family) *family = best;
  return MSH_OK;
}

int msh_autotune_batch(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                       int32_t* chosen) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (n < 0 || (n > 0 && (!fields || !params)))
    return fail(c, MSH_ERR_ARGUMENT, "null fields or params");
  // the groups that choose together: (method, element type, size class)
  std::vector<int> group((size_t)n);
  std::vector<std::vector<int>> members;
  std::vector<int> gkey;
  for (int v = 0; v < n; ++v) {
    Plan pl;
    int rc = make_plan(c, &fields[v], &params[v], &pl);
    if (rc) return rc;
    const int k = (pl.method * 3 + pl.T) * kSizeClasses + size_class(&fields[v]);
    size_t g = 0;
    while (g < gkey.size() && gkey[g] != k) ++g;
    if (g == gkey.size()) {
      gkey.push_back(k);
      members.emplace_back();
    }
    members[g].push_back(v);
    group[v] = (int)g;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int rc = e == hipSuccess ? MSH_OK : hip_fail(c, e, "autotune events");
  std::vector<int> best((size_t)gkey.size(), -1);
  for (size_t g = 0; g < gkey.size() && rc == MSH_OK; ++g) {
    const int method = gkey[g] / (3 * kSizeClasses), T = (gkey[g] / kSizeClasses) % 3,
              cls = gkey[g] % kSizeClasses;
    const size_t m = members[g].size();
    std::vector<msh_field> fs(m);
    std::vector<msh_params> ps(m);
    for (size_t i = 0; i < m; ++i) {
      fs[i] = fields[members[g][i]];
      ps[i] = params[members[g][i]];
    }
    float best_ms = 0.f;
    for (int fam = 3; fam >= 0 && rc == MSH_OK; --fam) {
      c->family[method][T][cls] = fam;
      std::vector<int64_t> nv(m), nf(m);
      rc = msh_batch_count(ctx, (int32_t)m, fs.data(), ps.data(), nv.data(), nf.data());
      if (rc) break;
      // every volume of the group must take this family (else it is the
      // same launch as a lower family's)
      bool takes = true;
      for (size_t i = 0; i < m && takes; ++i) {
        Plan q = c->batch.plans[i];
        VolArgs probe = c->batch.vols[i];
        takes = variant_family(emit_variant(q, &probe)) == fam;
      }
      if (!takes) continue;
      size_t total = 256;
      for (size_t i = 0; i < m; ++i)
        total += align256((size_t)nv[i] * 3 * (c->batch.plans[i].Out == K_F32 ? 4 : 8)) +
                 align256((size_t)nf[i] * 24);
      rc = grow(c, &c->ostage, &c->ostage_bytes, total, "autotune output buffer");
      if (rc) break;
      std::vector<void*> vp(m);
      std::vector<int64_t*> fp(m);
      size_t off = 0;
      for (size_t i = 0; i < m; ++i) {
        vp[i] = (char*)c->ostage + off;
        off += align256((size_t)nv[i] * 3 * (c->batch.plans[i].Out == K_F32 ? 4 : 8));
        fp[i] = (int64_t*)((char*)c->ostage + off);
        off += align256((size_t)nf[i] * 24);
      }
      float ms = 0.f;
      for (int rep = 0; rep < 4 && rc == MSH_OK; ++rep) {  // a warm-up, then the best of 3
        e = hipEventRecord(e0, c->stream);
        if (e == hipSuccess)
          rc = msh_batch_extract(ctx, (int32_t)m, fs.data(), ps.data(), vp.data(), fp.data(),
                                 MSH_DEVICE);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e != hipSuccess) rc = hip_fail(c, e, "autotune");
        if (rep == 1 || (rep > 1 && t < ms)) ms = t;
      }
      if (rc == MSH_OK && (best[g] < 0 || ms < best_ms)) {
        best[g] = fam;
        best_ms = ms;
      }
    }
    c->family[method][T][cls] = best[g];
  }
  // Then the whole batch: its groups' launches run side by side (Spread),
  // so a group's fastest shape alone need not be fastest beside the others.
  // One pass over the groups, most volumes first, keeps a change of one
  // group's family when it takes the whole batch's emit 1% below the best so
  // far (skipped when the device has no room for all the outputs at once).
  // took: every volume of group g takes family fam.
  bool no_room = false;
  auto time_all = [&](int g, int fam, float* ms, bool* took) -> int {
    std::vector<int64_t> nv((size_t)n), nf((size_t)n);
    int r = msh_batch_count(ctx, n, fields, params, nv.data(), nf.data());
    if (r) return r;
    *took = true;
    if (g >= 0)
      for (int v : members[(size_t)g]) {
        VolArgs probe = c->batch.vols[v];
        *took = *took && variant_family(emit_variant(c->batch.plans[v], &probe)) == fam;
      }
    if (!*took) return MSH_OK;
    size_t total = 256;
    for (int v = 0; v < n; ++v)
      total += align256((size_t)nv[v] * 3 * (c->batch.plans[v].Out == K_F32 ? 4 : 8)) +
               align256((size_t)nf[v] * 24);
    r = grow(c, &c->ostage, &c->ostage_bytes, total, "autotune output buffer");
    if (r) {  // no room for the whole batch's outputs: the group choices stand
      (void)hipGetLastError();
      c->err.clear();
      *took = false;
      no_room = true;
      return MSH_OK;
    }
    std::vector<void*> vp((size_t)n);
    std::vector<int64_t*> fp((size_t)n);
    size_t off = 0;
    for (int v = 0; v < n; ++v) {
      vp[v] = (char*)c->ostage + off;
      off += align256((size_t)nv[v] * 3 * (c->batch.plans[v].Out == K_F32 ? 4 : 8));
      fp[v] = (int64_t*)((char*)c->ostage + off);
      off += align256((size_t)nf[v] * 24);
    }
    for (int rep = 0; rep < 4; ++rep) {  // a warm-up, then the best of 3
      hipError_t he = hipEventRecord(e0, c->stream);
      if (he != hipSuccess) return hip_fail(c, he, "autotune");
      r = msh_batch_extract(ctx, n, fields, params, vp.data(), fp.data(), MSH_DEVICE);
      if (r) return r;
      he = hipEventRecord(e1, c->stream);
      if (he == hipSuccess) he = hipEventSynchronize(e1);
      float t = 0.f;
      if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
      if (he != hipSuccess) return hip_fail(c, he, "autotune");
      if (rep == 1 || (rep > 1 && t < *ms)) *ms = t;
    }
    return MSH_OK;
  };
  if (rc == MSH_OK && gkey.size() > 1) {
    float cur = 0.f;
    bool took = true;
    rc = time_all(-1, 0, &cur, &took);
    // (with no room for the whole batch there is nothing to compare against)
    std::vector<size_t> gord(rc == MSH_OK && !no_room ? gkey.size() : 0);
    for (size_t g = 0; g < gord.size(); ++g) gord[g] = g;
    std::stable_sort(gord.begin(), gord.end(),
                     [&](size_t x, size_t y) { return members[x].size() > members[y].size(); });
    for (size_t g : gord) {
      const int method = gkey[g] / (3 * kSizeClasses), T = (gkey[g] / kSizeClasses) % 3,
                cls = gkey[g] % kSizeClasses;
      for (int fam = 3; fam >= 0 && rc == MSH_OK; --fam) {
        if (fam == best[g]) continue;
        c->family[method][T][cls] = fam;
        float ms = 0.f;
        rc = time_all((int)g, fam, &ms, &took);
        if (no_room) break;
        if (rc == MSH_OK && took && ms < 0.99f * cur) {
          best[g] = fam;
          cur = ms;
        }
      }
      c->family[method][T][cls] = best[g];
      if (rc || no_room) break;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  c->batch.have = false;  // the timing batches are not the caller's
  c->have = false;
  if (rc) return rc;
  if (chosen)
    for (int v = 0; v < n; ++v) chosen[v] = best[(size_t)group[v]];
  return MSH_OK;
}

int msh_weld(void* ctx, int32_t kind, int64_t nverts, const void* verts, int64_t nfaces,
             int64_t* faces, void* verts_out, int64_t* nunique) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (kind != K_F32 && kind != K_F64)
    return fail(c, MSH_ERR_UNSUPPORTED, "vertex coordinates must be Float32 or Float64");
  if (nverts < 0 || nfaces < 0) return fail(c, MSH_ERR_ARGUMENT, "negative size");
  if (!nunique) return fail(c, MSH_ERR_ARGUMENT, "null nunique");
  if ((nverts > 0 && (!verts || !verts_out)) || (nfaces > 0 && !faces))
    return fail(c, MSH_ERR_ARGUMENT, "null buffer");
  const size_t vbytes = (size_t)nverts * 3 * (kind == K_F32 ? 4 : 8);
  if (nverts > 0 && (const char*)verts_out < (const char*)verts + vbytes &&
      (const char*)verts < (const char*)verts_out + vbytes)
    return fail(c, MSH_ERR_ARGUMENT, "verts_out overlaps verts");
  // hash slots are noted as 32-bit numbers: at most 2^32 slots
  if (nverts > ((int64_t)3 << 30) || nfaces > (int64_t)1 << 40)
    return fail(c, MSH_ERR_ARGUMENT, "mesh too large (at most 3 * 2^30 vertices)");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  // workspace: flags | prefixes | first appearances | tiles | sizes | region bounds | hash table
  int64_t nslots = 1;
  while (nslots < nverts + nverts / 4 + 1) nslots <<= 1;
  const int64_t U = nverts > 0 ? nverts : 1;
  const int64_t T = (U + 4095) / 4096;
  const size_t b_flag = align256((size_t)U * 4), b_i64 = align256((size_t)U * 8);
  const size_t b_tcnt = align256((size_t)(T + 1) * 4), b_toff = align256((size_t)T * 8);
  const size_t b_tot = 256, b_bnd = 512, b_table = align256((size_t)nslots * 8);
  const size_t need = b_flag + 2 * b_i64 + b_tcnt + b_toff + b_tot + b_bnd + b_table;
  int rc = grow(c, &c->wws, &c->wws_bytes, need, "weld workspace allocation");
  if (rc) return rc;
  WeldArgs a{};
  char* q = (char*)c->wws;
  a.flag = (uint32_t*)q;
  q += b_flag;
  a.pos = (int64_t*)q;
  q += b_i64;
  a.rep = (int64_t*)q;
  q += b_i64;
  a.tilecnt = (uint32_t*)q;
  q += b_tcnt;
  a.tileoff = (int64_t*)q;
  q += b_toff;
  int64_t* sizes = (int64_t*)q;  // [0] welded count, [1] bad face indices
  a.total = sizes;
  a.bad = sizes + 1;
  q += b_tot;
  a.bounds = (uint64_t*)q;
  q += b_bnd;
  a.table = (int64_t*)q;
  a.out_kind = kind;
  a.nverts = nverts;
  a.nfaces = nfaces;
  a.verts = verts;
  a.verts_out = verts_out;
  a.faces = faces;
  a.nslots = nslots;
  hipError_t e = launch_weld(a, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->pinned, sizes, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return hip_fail(c, e, "weld");
  const int64_t bad = c->pinned[1];
  if (bad != 0)
    return fail(c, MSH_ERR_ARGUMENT,
                std::to_string(bad) + " face indices outside 1.." + std::to_string(nverts));
  *nunique = nverts > 0 ? c->pinned[0] : 0;
  return MSH_OK;
}

int msh_generate_field(void* ctx, int32_t kind, int32_t dtype, int64_t nx, int64_t ny,
                       int64_t nz, uint64_t seed, double p0, double p1, void* out) {
  Ctx* c = (Ctx*)ctx;
  if (!c) return MSH_ERR_ARGUMENT;
  if (kind < 0 || kind > 4) return fail(c, MSH_ERR_ARGUMENT, "unknown field kind");
  if (dtype != K_F32 && dtype != K_F64) return fail(c, MSH_ERR_UNSUPPORTED, "dtype");
  if (nx < 0 || ny < 0 || nz < 0) return fail(c, MSH_ERR_ARGUMENT, "negative size");
  if (nx * ny * nz > 0 && !out) return fail(c, MSH_ERR_ARGUMENT, "null output");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, MSH_ERR_DEVICE, "hipSetDevice");
  hipError_t e = launch_generate(kind, dtype, nx, ny, nz, seed, p0, p1, out, c->stream);
  return e == hipSuccess ? MSH_OK : hip_fail(c, e, "generate");
}

}  // extern "C"
