/* c_abi_check.c -- a caller of libmeshing_hip.so through its C ABI alone
 * (include/meshing_hip.h), the way the Julia ccall binding would use it: no
 * Python, no torch.  Meshes a host Float32 field with each method (count,
 * allocate exactly, extract into host memory), and a batch of the three,
 * and writes every mesh to a binary file for the test to compare with the
 * CPU oracle:
 *   per mesh: int64 nverts, int64 nfaces, int32 vertex kind (1 Float32,
 *   2 Float64), then the vertices, then the faces (int64, 1-based).
 * usage: c_abi_check NX NY NZ FIELD.f32 OUT.bin
 * FIELD.f32 holds NX*NY*NZ float32 values, x fastest. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/meshing_hip.h"

static void* ctx_;

static int check(int st, const char* what) {
  if (st != MSH_OK) {
    fprintf(stderr, "%s: status %d: %s\n", what, st, ctx_ ? msh_last_error(ctx_) : "");
    exit(1);
  }
  return st;
}

static msh_params params(int method) {
  msh_params p;
  memset(&p, 0, sizeof p);
  p.method = method;
  p.iso_kind = MSH_INT64; /* iso = 0 (Julia Int) */
  p.iso_int = 0;
  p.eps_kind = MSH_FLOAT64; /* MarchingTetrahedra's default eps = 1e-3 */
  p.eps = 1e-3;
  for (int k = 0; k < 3; ++k) {
    p.axis_kind[k] = MSH_INT64; /* X = Y = Z = -1:1 */
    p.first[k] = -1.0;
    p.last[k] = 1.0;
  }
  return p;
}

static void write_mesh(FILE* out, int64_t nv, int64_t nf, int32_t kind, const void* v,
                       const int64_t* f) {
  fwrite(&nv, 8, 1, out);
  fwrite(&nf, 8, 1, out);
  fwrite(&kind, 4, 1, out);
  fwrite(v, kind == MSH_FLOAT32 ? 4 : 8, (size_t)(3 * nv), out);
  fwrite(f, 8, (size_t)(3 * nf), out);
}

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s NX NY NZ FIELD.f32 OUT.bin\n", argv[0]);
    return 2;
  }
  const int64_t nx = atoll(argv[1]), ny = atoll(argv[2]), nz = atoll(argv[3]);
  const size_t n = (size_t)(nx * ny * nz);
  float* sdf = (float*)malloc(n * sizeof(float));
  FILE* in = fopen(argv[4], "rb");
  if (!sdf || !in || fread(sdf, sizeof(float), n, in) != n) {
    fprintf(stderr, "cannot read the field\n");
    return 2;
  }
  fclose(in);
  FILE* out = fopen(argv[5], "wb");
  if (!out) return 2;
  check(msh_create(0, &ctx_), "msh_create");
  const msh_field f = {MSH_FLOAT32, MSH_HOST, nx, ny, nz, sdf};
  msh_field fs[3];
  msh_params ps[3];
  for (int method = 0; method < 3; ++method) {
    const msh_params p = params(method);
    int32_t kind = 0;
    int64_t nv = 0, nf = 0;
    check(msh_output_kind(&f, &p, &kind), "msh_output_kind");
    check(msh_count(ctx_, &f, &p, &nv, &nf), "msh_count");
    void* v = malloc((size_t)(3 * nv) * (kind == MSH_FLOAT32 ? 4 : 8) + 8);
    int64_t* fc = (int64_t*)malloc((size_t)(3 * nf) * 8 + 8);
    check(msh_extract(ctx_, &f, &p, v, fc, MSH_HOST), "msh_extract");
    write_mesh(out, nv, nf, kind, v, fc);
    printf("method %d: %lld vertices, %lld faces\n", method, (long long)nv, (long long)nf);
    free(v);
    free(fc);
    fs[method] = f;
    ps[method] = p;
  }
  /* the same three as one batch */
  int64_t nv[3], nf[3];
  void* vv[3];
  int64_t* ff[3];
  int32_t kinds[3];
  check(msh_batch_count(ctx_, 3, fs, ps, nv, nf), "msh_batch_count");
  for (int v = 0; v < 3; ++v) {
    check(msh_output_kind(&fs[v], &ps[v], &kinds[v]), "msh_output_kind");
    vv[v] = malloc((size_t)(3 * nv[v]) * (kinds[v] == MSH_FLOAT32 ? 4 : 8) + 8);
    ff[v] = (int64_t*)malloc((size_t)(3 * nf[v]) * 8 + 8);
  }
  check(msh_batch_extract(ctx_, 3, fs, ps, vv, ff, MSH_HOST), "msh_batch_extract");
  for (int v = 0; v < 3; ++v) {
    write_mesh(out, nv[v], nf[v], kinds[v], vv[v], ff[v]);
    free(vv[v]);
    free(ff[v]);
  }
  /* errors come back as statuses with a message */
  msh_field bad = f;
  bad.dtype = 7;
  const msh_params p0 = params(0);
  int64_t a, b;
  const int st = msh_count(ctx_, &bad, &p0, &a, &b);
  printf("bad dtype: status %d (%s)\n", st, msh_last_error(ctx_));
  check(msh_destroy(ctx_), "msh_destroy");
  fclose(out);
  free(sdf);
  return st == MSH_ERR_UNSUPPORTED ? 0 : 1;
}
