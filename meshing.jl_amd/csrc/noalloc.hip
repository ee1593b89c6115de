// noalloc.hip -- the device scratch of msh_isosurface_noalloc, the extraction
// entry point that allocates nothing.
//
// The count/scan workspace of that entry point (per-unit counts, z-block
// masks and output prefixes, the scan's tile sums and offsets, the totals,
// the coordinate tables)
// lives in one file-scope device array of kScratchBytes.  Every GPU the
// library runs on holds its own copy of the array, at its own address, for
// the life of the process.  Host code never passes the array's name to a
// kernel (in host code that name is a stand-in, not a device address): at
// context creation noalloc_setup() looks up this device's address of the
// array once, with the device current, and kernels get that pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

namespace msh {
namespace {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The clamped sizes the emit kernels read (first vertex, vertices to write,
// first face, faces to write) and the exact sizes for the caller (null
// totals: a volume without cells).
__global__ void k_noalloc_sizes(const int64_t* __restrict__ totals, int64_t vcap, int64_t fcap,
                                int64_t* __restrict__ clamp, int64_t* __restrict__ sizes) {
  if (threadIdx.x == 0) {
    const int64_t nv = totals ? totals[1] : 0, nf = totals ? totals[3] : 0;
    clamp[0] = 0;
    clamp[1] = nv < vcap ? nv : vcap;
    clamp[2] = 0;
    clamp[3] = nf < fcap ? nf : fcap;
    if (sizes) {
      sizes[0] = nv;
      sizes[1] = nf;
    }
  }
}

}  // namespace

hipError_t noalloc_setup(int device, void** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    // the caller (msh_create) has made `device` current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
  }
  *ptr = g_ptr[device];
  return hipSuccess;
}

hipError_t launch_noalloc_sizes(const int64_t* totals, int64_t vcap, int64_t fcap, int64_t* clamp,
                                int64_t* sizes, hipStream_t st) {
  hipLaunchKernelGGL(k_noalloc_sizes, dim3(1), dim3(64), 0, st, totals, vcap, fcap, clamp, sizes);
  return hipGetLastError();
}

}  // namespace msh
