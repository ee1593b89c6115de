/* meshing_hip.h -- C ABI of libmeshing_hip.so, MI355X (gfx950) isosurface
 * extraction: Marching Cubes, Marching Tetrahedra, Naive Surface Nets.
 *
 * Drop-in boundary: the reference's generic function
 *     isosurface(sdf::AbstractArray{T,3}, method, X=-1:1, Y=-1:1, Z=-1:1)
 *         -> (Vector{NTuple{3,float(FT)}}, Vector{NTuple{3,Int}})
 * (Meshing.jl v0.7.0 src/isosurface.jl:30-32, src/marching_cubes.jl:27-64,
 * src/marching_tetrahedra.jl:129-163, src/algorithmtypes.jl:10,23-25,37-40).
 * A Julia `ccall` (or Python ctypes) wrapper computes first(X)/last(X) and
 * the Julia types of iso/eps/coordinates, calls msh_count to learn the output
 * sizes, allocates, and calls msh_extract.  Results are identical to the
 * reference's: same vertex and face order, 1-based Int64 faces, vertex
 * coordinates rounded to float(FT) after the same mixed-precision arithmetic.
 *
 * All functions return an msh_status; on error msh_last_error(ctx) holds a
 * message.  The library never aborts.  A context owns one HIP stream (or
 * uses the caller's, see msh_set_stream) and grow-only device workspaces;
 * a context must not be used by two threads at once.  Separate contexts are
 * independent.
 */
#ifndef MESHING_HIP_H
#define MESHING_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  MSH_OK = 0,
  MSH_ERR_ARGUMENT = 1,    /* Julia ArgumentError / MethodError analogue */
  MSH_ERR_OUT_OF_MEMORY = 2,
  MSH_ERR_DEVICE = 3,      /* HIP runtime failure */
  MSH_ERR_STATE = 4,       /* msh_extract without a matching msh_count */
  MSH_ERR_UNSUPPORTED = 5  /* element type the GPU path does not handle */
} msh_status;

/* AbstractMeshingAlgorithm subtypes (src/algorithmtypes.jl:10). */
typedef enum {
  MSH_MARCHING_CUBES = 0,       /* MarchingCubes{T}(iso)           :23-25 */
  MSH_MARCHING_TETRAHEDRA = 1,  /* MarchingTetrahedra{T,E}(iso,eps) :37-40 */
  MSH_NAIVE_SURFACE_NETS = 2    /* named by the reference, no code in v0.7.0 */
} msh_method;

/* Julia scalar types that take part in promotion. */
typedef enum { MSH_INT64 = 0, MSH_FLOAT32 = 1, MSH_FLOAT64 = 2 } msh_kind;

/* Where a buffer lives. */
typedef enum { MSH_HOST = 0, MSH_DEVICE = 1 } msh_memory;

/* The sampled field `sdf`: column-major (x fastest), size(sdf) = (nx,ny,nz). */
typedef struct {
  int32_t dtype;   /* eltype(sdf): MSH_FLOAT32 or MSH_FLOAT64 */
  int32_t memory;  /* MSH_HOST or MSH_DEVICE */
  int64_t nx, ny, nz;
  const void* data;
} msh_field;

/* The method object and the X, Y, Z axis arguments. */
typedef struct {
  int32_t method;              /* msh_method */
  int32_t iso_kind;            /* typeof(method.iso) */
  int32_t eps_kind;            /* typeof(method.eps), MT only */
  int32_t axis_kind[3];        /* eltype(first(X)), eltype(first(Y)), eltype(first(Z)) */
  double iso, eps;             /* value when the kind is a float */
  int64_t iso_int, eps_int;    /* value when the kind is MSH_INT64 */
  double first[3], last[3];    /* first(X), last(X), ... (Int endpoints as Float64) */
} msh_params;

/* Context life cycle.  `device` is a HIP device ordinal.  The first context
 * on a device also builds that device's lookup-table image (one short kernel
 * on the new context's stream, waited for before msh_create returns). */
int msh_create(int device, void** ctx);
int msh_destroy(void* ctx);
/* Run subsequent work on the caller's hipStream_t; NULL is HIP's default
 * (null) stream.  A new context uses a non-blocking stream of its own. */
int msh_set_stream(void* ctx, void* hip_stream);
int msh_synchronize(void* ctx);
const char* msh_last_error(const void* ctx);

/* float(FT) of the result (src/marching_cubes.jl:31,33; src/marching_tetrahedra.jl:131,135):
 * MSH_FLOAT32 or MSH_FLOAT64. */
int msh_output_kind(const msh_field* field, const msh_params* params, int32_t* kind);

/* Phase 1 of isosurface(): classify every cell, count, and scan the counts
 * in the reference's emission order (x outermost, z innermost).  Returns the
 * exact output sizes.  The scan stays in the context for msh_extract. */
int msh_count(void* ctx, const msh_field* field, const msh_params* params,
              int64_t* nverts, int64_t* nfaces);

/* Phase 2: write vertices (3*nverts values of msh_output_kind) and faces
 * (3*nfaces int64, 1-based) for the field/params of the preceding msh_count
 * on this context.  `out_memory` says where verts/faces live.  Device outputs
 * are written asynchronously on the context's stream; host outputs are
 * complete on return. */
int msh_extract(void* ctx, const msh_field* field, const msh_params* params,
                void* verts, int64_t* faces, int32_t out_memory);

/* Batched form: n independent volumes (any sizes, element types, methods
 * and parameters) through one classify-count-scan sequence and one emit
 * launch per kernel instantiation, instead of n separate calls.  Element v of
 * the results is what msh_count / msh_extract return for fields[v] and
 * params[v] alone (same reference functions as those two).  msh_batch_count
 * writes nverts[v] and nfaces[v]; msh_batch_extract, called next with the same
 * arrays, writes volume v into verts[v] and faces[v] (faces 1-based within
 * the volume).  Fields may be host or device memory; outputs as out_memory
 * says.  A batch shares the context's workspace with the single-volume calls:
 * msh_count in between invalidates it (MSH_ERR_STATE).  The count and emit
 * launches of a batch spread over the context's stream and up to two more
 * that the context creates (default priority), forked from and joined back
 * to its stream by events, so to the caller the batch stays ordered on that
 * one stream. */
int msh_batch_count(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                    int64_t* nverts, int64_t* nfaces);
int msh_batch_extract(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                      void* const* verts, int64_t* const* faces, int32_t out_memory);
/* A batch whose arguments repeat the previous call's (same fields, params,
 * kernel choices and workspace buffers, device fields, device outputs, a
 * stream other than the null stream) runs each phase from a hipGraph: the
 * second such call captures the phase's copies and launches, later calls
 * replay it (new output pointers travel through the descriptor copy, so
 * they may change).  The count and emit kernels of the batch are persistent:
 * as many blocks as fit on the device take the batch's tiles from a work
 * queue.  msh_batch_graph_launches reports how many times each phase has run
 * from a graph on this context.  No counterpart in the reference. */
int msh_batch_graph_launches(void* ctx, int64_t* count_phase, int64_t* emit_phase);

/* Kernel choice.  The emit pass has kernel families per method: 2 lanes
 * along x (the default), 1 lanes along z, 0 a generic kernel, and 3 the
 * lanes-along-x kernel with half-depth bricks (a tile shape, for Naive
 * Surface Nets and for Marching Cubes with one element type throughout and
 * Float64 axes; otherwise the same as 2); each family falls back to the
 * next one down for volumes it cannot take.
 * msh_autotune counts the given field, times the emit of every family that
 * takes it (the best of 3 to 50 launches each, about 2 ms, on the
 * context's stream, into a scratch buffer) and records the fastest for the
 * field's (method, element type, size class) in the context, where later
 * msh_count and msh_batch_count calls on volumes of that class use it (size
 * classes by cell count: below 2^20, below 2^23, larger; a batch of mixed
 * sizes then runs each volume with its class's choice);
 * *family returns the choice.  msh_set_kernel_family sets it directly for
 * every size class (-1 restores the default).  Results do not depend on the
 * family.  No counterpart in the reference (its tile shapes are the CPU
 * loop). */
int msh_autotune(void* ctx, const msh_field* field, const msh_params* params, int32_t* family);
/* msh_autotune for a batch: the batch's volumes are grouped by (method,
 * element type, size class), and each group's batched emit (all its volumes
 * in one msh_batch_extract, device outputs in a scratch buffer) is timed
 * with every family that takes all of them (a warm-up, then the best of 3);
 * the fastest is recorded for the group; then, since the groups' launches
 * run side by side, one pass over the groups (most volumes first) times the
 * whole batch with each other family that takes the group and keeps a
 * change that makes it 1% faster.  chosen[v] (n entries, may be null)
 * returns volume v's.  A batch runs the persistent batched kernels,
 * whose best shape differs from a lone small volume's (which cannot fill
 * the device by itself).  Leaves no batch state behind: call
 * msh_batch_count next. */
int msh_autotune_batch(void* ctx, int32_t n, const msh_field* fields, const msh_params* params,
                       int32_t* chosen);
int msh_set_kernel_family(void* ctx, int32_t method, int32_t dtype, int32_t family);

/* Vertex welding, the opt-in step after extraction: vertices whose three
 * coordinates have identical bit patterns merge into their first
 * appearance, the welded vertices keep first-appearance order, and the
 * faces (1-based) are renumbered in place.  verts holds nverts x 3
 * coordinates of `kind` (MSH_FLOAT32 or MSH_FLOAT64); verts, faces and
 * verts_out are device memory, verts_out with room for nverts x 3 and
 * apart from verts; it receives the *nunique welded vertices.  A face index outside 1..nverts
 * fails with MSH_ERR_ARGUMENT (the valid ones are renumbered by then), as
 * do meshes of more than 3 * 2^30 vertices.  Returns once *nunique is
 * known.  The reference's Marching Cubes repeats
 * the vertices that neighbouring cubes share (src/algorithmtypes.jl:17-19)
 * and leaves merging them to the caller; it has no welding function. */
int msh_weld(void* ctx, int32_t kind, int64_t nverts, const void* verts, int64_t nfaces,
             int64_t* faces, void* verts_out, int64_t* nunique);

/* Synthetic fields, generated on the device into a column-major buffer of
 * nx*ny*nz elements of `dtype`.  kind: 0 sphere sqrt(x^2+y^2+z^2)-p0 on
 * [-1,1]^3; 1 torus (R=p0, r=p1) on [-1,1]^3; 2 gyroid on [-p0,p0]^3;
 * 3 value noise (32^3 lattice + 0.25 x 128^3 octave, seeded); 4 sum of 16
 * seeded Gaussians on [-1,1]^3.  No counterpart in the reference (its tests
 * build fields with comprehensions, test/runtests.jl:10-13). */
int msh_generate_field(void* ctx, int32_t kind, int32_t dtype, int64_t nx, int64_t ny,
                       int64_t nz, uint64_t seed, double p0, double p1, void* out_device);

/* Extraction that allocates nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is
 * msh_count followed by msh_extract in one call, on a device field into
 * caller-owned device outputs with room for vcap vertices and fcap faces;
 * vertices beyond vcap and faces beyond fcap are not written.  The call
 * queues its work on the context's stream and returns without waiting; when
 * that work has run, sizes[0] and sizes[1] (device memory) hold the exact
 * vertex and face counts, so a caller whose capacity was short sees it there.
 * The count/scan workspace is a static device array of
 * MSH_NOALLOC_SCRATCH_BYTES per GPU, shared by every call on that GPU and
 * holding the last call's contents: calls that overlap on one GPU (two
 * contexts, or two streams) race on it and must be serialised by the caller.
 * A volume whose workspace does not fit the array (about 28 bytes per z row
 * of cells, each row taken whole up to 4096 cells, so up to about 1.2 M
 * rows, e.g. 1024 x 1024 x nz for nz up to 4097) fails
 * with MSH_ERR_ARGUMENT, as do host fields.  Results equal msh_extract's. */
#define MSH_NOALLOC_SCRATCH_BYTES (32u << 20)
int msh_isosurface_noalloc(void* ctx, const msh_field* field, const msh_params* params,
                           void* verts, int64_t vcap, int64_t* faces, int64_t fcap,
                           int64_t* sizes);

#ifdef __cplusplus
}
#endif

#endif /* MESHING_HIP_H */
