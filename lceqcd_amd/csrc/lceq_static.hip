// lceq_static.hip -- lceq_check_fields_static: the field guard of
// lceq_check_fields with its per-block partial sums and their reduction in a
// fixed device array instead of memory the library allocates, for callers
// that cannot allocate device memory (code under stream capture, serving
// loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
#include <mutex>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"

using namespace lceq;

// largest number of 256-site blocks (2^24 sites per parity) whose partial
// sums the fixed array holds: 1 MiB of device memory per GPU
#define LCEQ_STATIC_MAX_BLOCKS 65536

namespace {

// per-block partials of k_check_fields (2 per 256-site block) and the two
// reduced values after them
constexpr int SCRATCH_DOUBLES = 2 * LCEQ_STATIC_MAX_BLOCKS + 2;
constexpr int MAX_DEVICES = 64;

// One copy per device, held until the process ends.  In host code this name
// is not a device address: only the pointer hipGetSymbolAddress returns for a
// device (scratch_ptr) is used there.
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

std::mutex g_mu;
double* g_ptr[MAX_DEVICES];

// the device address of `scratch` on `device`, looked up once per device
// with that device current
int scratch_ptr(int device, double** out) {
  if (device < 0 || device >= MAX_DEVICES)
    return set_error(LCEQ_ERR_ARG, "device %d beyond the %d supported", device, MAX_DEVICES);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return set_error(LCEQ_ERR_HIP, "hipGetDevice failed");
    if (cur != device && hipSetDevice(device) != hipSuccess)
      return set_error(LCEQ_ERR_HIP, "hipSetDevice(%d) failed", device);
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p)
      return set_error(LCEQ_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_ptr[device] = static_cast<double*>(p);
  }
  *out = g_ptr[device];
  return LCEQ_OK;
}

// One block: the largest deviation and the sum of the non-finite counts over
// the nb partials (max is order-free; the counts are integers, exact in any
// order), into out[0], out[1].
__global__ __launch_bounds__(256) void k_check_reduce(const double* __restrict__ part, int nb,
                                                      double* __restrict__ out) {
  __shared__ double wmax[4], wbad[4];
  double dev = 0.0, bad = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) {
    dev = fmax(dev, part[2 * b]);
    bad += part[2 * b + 1];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    dev = fmax(dev, __shfl_xor(dev, off, 64));
    bad += __shfl_xor(bad, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    wmax[threadIdx.x >> 6] = dev;
    wbad[threadIdx.x >> 6] = bad;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    out[1] = ((wbad[0] + wbad[1]) + wbad[2]) + wbad[3];
  }
}

}  // namespace

extern "C" int lceq_check_fields_static(lceq_lattice* L, double tol, double* max_dev,
                                        int64_t* nonfinite) {
  if (!L) return set_error(LCEQ_ERR_ARG, "null lattice handle");
  int nb = blocks_for(L->g.n);
  if (nb > LCEQ_STATIC_MAX_BLOCKS)
    return set_error(LCEQ_ERR_ARG,
                     "lceq_check_fields_static: %d sites per parity need %d partial sums, above "
                     "the fixed maximum of %d (%d sites per parity)",
                     L->g.n, nb, LCEQ_STATIC_MAX_BLOCKS, 256 * LCEQ_STATIC_MAX_BLOCKS);
  double* ptr = nullptr;
  int rc = scratch_ptr(L->device, &ptr);
  if (rc) return rc;
  // every partial the reduction reads is written by this call's first
  // launch, so nothing left by an earlier call needs clearing
  launch_check_fields(L, ptr);
  double* res = ptr + 2 * LCEQ_STATIC_MAX_BLOCKS;
  hipLaunchKernelGGL(k_check_reduce, dim3(1), dim3(256), 0, L->stream, ptr, nb, res);
  if ((rc = check_launch())) return rc;
  double h[2];
  hipError_t e = hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, L->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L->stream);
  if (e != hipSuccess)
    return set_error(LCEQ_ERR_HIP, "reading the field guard: %s", hipGetErrorString(e));
  if (max_dev) *max_dev = h[0];
  if (nonfinite) *nonfinite = (int64_t)h[1];
  if (h[1] > 0.0) return set_error(LCEQ_ERR_FIELD, "%.0f non-finite link or Higgs values", h[1]);
  if (tol >= 0.0 && h[0] > tol)
    return set_error(LCEQ_ERR_FIELD, "links off SU(2) by %.3g (tolerance %.3g)", h[0], tol);
  return LCEQ_OK;
}
