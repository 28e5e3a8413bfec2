// lceq_api.hip -- C ABI of liblceq.so (declared in include/lceq.h): lattice
// lifecycle, host <-> device field exchange in the reference COMMON layout,
// and one entry point per reference routine of the update and measurement
// path.  Every entry point queues kernels on the lattice's stream; only the
// ones that return host values synchronise.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <string>
#include <vector>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"

using namespace lceq;

static thread_local std::string g_err;

int lceq::set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
#define fail lceq::set_error

#define HIPCHK(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(LCEQ_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
  } while (0)

#define NEEDL(L) \
  if (!(L)) return fail(LCEQ_ERR_ARG, "null lattice handle")

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LCEQ_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  return LCEQ_OK;
}

extern "C" const char* lceq_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------------ lifecycle

extern "C" int lceq_create(int lx, int ly, int lz, lceq_lattice** out) {
  if (!out) return fail(LCEQ_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (const char* why = geom_reject(lx, ly, lz))
    return fail(LCEQ_ERR_ARG, "lattice %d x %d x %d: %s", lx, ly, lz, why);
  lceq_lattice* L = new lceq_lattice();
  memset(L, 0, sizeof(*L));
  L->g = geom_init(lx, ly, lz);
  Geom& g = L->g;
  L->vol = lx * ly * lz;
  L->tb = 256;
  L->xcd = 1;
  L->pstride = (g.n + 63) / 64 + 8;  // block partials of the smallest launch shape
  L->beta = 9.0;
  L->lambda = 1.0;
  L->alpha = 1.0;
  L->limit_or = 1e-9;
  L->max_iter = 1000000;
  L->rng_mode = LCEQ_RNG_PHILOX;
  L->seed = 12345;
  int rc = LCEQ_OK;
  auto alloc = [&](double** p, size_t count) {
    if (rc != LCEQ_OK) return;
    hipError_t e = hipMalloc((void**)p, count * sizeof(double));
    if (e != hipSuccess) rc = fail(LCEQ_ERR_HIP, "hipMalloc(%zu doubles): %s", count, hipGetErrorString(e));
  };
  size_t np = (size_t)g.npad;
  if (hipGetDevice(&L->device) != hipSuccess) rc = fail(LCEQ_ERR_HIP, "no HIP device");
  if (rc == LCEQ_OK && hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking) != hipSuccess)
    rc = fail(LCEQ_ERR_HIP, "hipStreamCreate failed");

  alloc(&L->U, 24 * np);
  alloc(&L->A, 6 * np);
  alloc(&L->Uf, 24 * np);
  alloc(&L->Af, 6 * np);
  alloc(&L->Vst, 24 * np);
  alloc(&L->F, 3 * np);
  alloc(&L->partials, (size_t)LCEQ_NSLOT * LCEQ_NQ * L->pstride);

  alloc(&L->obs, LCEQ_OBS_N);
  alloc(&L->slices, 14 * (size_t)lz * corr_split(L->g));
  alloc(&L->corr, 5 * (size_t)lz + 8);
  L->hist_cap = 4096;
  alloc(&L->hist, (size_t)L->hist_cap * LCEQ_OBS_N);
  if (rc == LCEQ_OK) {
    // zero every field so padding lanes hold finite values
    (void)hipMemsetAsync(L->U, 0, 24 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->A, 0, 6 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->Uf, 0, 24 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->Af, 0, 6 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->Vst, 0, 24 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->F, 0, 3 * np * sizeof(double), L->stream);
    (void)hipMemsetAsync(L->partials, 0, (size_t)LCEQ_NSLOT * LCEQ_NQ * L->pstride * sizeof(double),
                   L->stream);
    (void)hipMemsetAsync(L->obs, 0, LCEQ_OBS_N * sizeof(double), L->stream);

    if (hipStreamSynchronize(L->stream) != hipSuccess) rc = fail(LCEQ_ERR_HIP, "init sync failed");
  }
  if (rc != LCEQ_OK) {
    lceq_destroy(L);
    return rc;
  }
  *out = L;
  return LCEQ_OK;
}

extern "C" int lceq_destroy(lceq_lattice* L) {
  if (!L) return LCEQ_OK;
  if (L->stream) (void)hipStreamSynchronize(L->stream);
  double* bufs[] = {L->U, L->A, L->Uf, L->Af, L->Vst, L->F, L->partials, L->obs,
                    L->slices, L->corr, L->hist, L->scratch, L->fed};
  for (double* b : bufs)
    if (b) (void)hipFree(b);
  if (L->prof_ev) {
    for (int i = 0; i < L->prof_cap; ++i) (void)hipEventDestroy(L->prof_ev[i]);
    delete[] L->prof_ev;
  }
  if (L->stream) (void)hipStreamDestroy(L->stream);
  delete L;
  return LCEQ_OK;
}

extern "C" int lceq_sync(lceq_lattice* L) {
  NEEDL(L);
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

// ----------------------------------------------------------- parameters

extern "C" int lceq_set_couplings(lceq_lattice* L, double beta, double beta_a, double beta_2,
                                  double beta_4) {
  NEEDL(L);
  L->beta = beta;
  L->beta_a = beta_a;
  L->beta_2 = beta_2;
  L->beta_4 = beta_4;
  return LCEQ_OK;
}

extern "C" int lceq_set_gauge_fixing(lceq_lattice* L, double lambda, double alpha,
                                     double limit_or, int max_iter) {
  NEEDL(L);
  if (max_iter < 1) return fail(LCEQ_ERR_ARG, "max_iter must be >= 1");
  L->lambda = lambda;
  L->alpha = alpha;
  L->limit_or = limit_or;
  L->max_iter = max_iter;
  return LCEQ_OK;
}

extern "C" int lceq_dims(lceq_lattice* L, int* dims) {
  NEEDL(L);
  if (!dims) return fail(LCEQ_ERR_ARG, "null output");
  dims[0] = L->g.lx;
  dims[1] = L->g.ly;
  dims[2] = L->g.lz;
  return LCEQ_OK;
}

extern "C" int lceq_set_shape(lceq_lattice* L, int threads_per_block, int xcd_order) {
  NEEDL(L);
  int t = threads_per_block;
  if (t != 64 && t != 128 && t != 256 && t != 512)
    return fail(LCEQ_ERR_ARG, "threads per block must be 64, 128, 256 or 512");
  HIPCHK(hipStreamSynchronize(L->stream));
  L->tb = t;
  L->xcd = xcd_order ? 1 : 0;
  return LCEQ_OK;
}

extern "C" int lceq_get_shape(lceq_lattice* L, int* threads_per_block, int* xcd_order) {
  NEEDL(L);
  if (threads_per_block) *threads_per_block = L->tb;
  if (xcd_order) *xcd_order = L->xcd;
  return LCEQ_OK;
}

// Plain device copy, one 16-byte element per thread
__global__ __launch_bounds__(256) void k_copy16(const double2* __restrict__ a,
                                                double2* __restrict__ b, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = a[i];
}

extern "C" int lceq_copy_bandwidth(int64_t bytes, int reps, double* us_per_copy) {
  if (!us_per_copy) return fail(LCEQ_ERR_ARG, "null output pointer");
  if (bytes < 16 || bytes % 16 || reps < 1 || (bytes / 16 + 255) / 256 > INT32_MAX)
    return fail(LCEQ_ERR_ARG, "copy size must be a positive multiple of 16 below 2^43, reps >= 1");
  int64_t n = bytes / 16;
  int nb = (int)((n + 255) / 256);
  double2 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LCEQ_OK;
  float el = 0.0f;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
      hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = fail(LCEQ_ERR_HIP, "copy buffers or events could not be set up");
  } else {
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy16, dim3(nb), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess)
      rc = fail(LCEQ_ERR_HIP, "copy timing failed");
    else
      *us_per_copy = 1e3 * (double)el / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// S streams read at the same index (16 bytes each per thread), one 16-byte
// write: the access shape of a launch that gathers S operand arrays
constexpr int STREAM_MAX = 32;
struct StreamPtrs {
  const double2* p[STREAM_MAX];
};
template <int S>
__global__ __launch_bounds__(256) void k_streams(StreamPtrs a, double2* __restrict__ out,
                                                 int64_t n) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double x = 0.0, y = 0.0;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    double2 v = a.p[k][i];
    x += v.x;
    y += v.y;
  }
  out[i] = double2{x, y};
}

extern "C" int lceq_stream_bandwidth(int streams, int64_t elements, int reps,
                                     double* us_per_pass) {
  if (!us_per_pass) return fail(LCEQ_ERR_ARG, "null output pointer");
  if (streams != 8 && streams != 16 && streams != 30)
    return fail(LCEQ_ERR_ARG, "streams must be 8, 16 or 30");
  if (elements < 1 || reps < 1 || (elements + 255) / 256 > INT32_MAX)
    return fail(LCEQ_ERR_ARG, "elements must be positive and below 2^39, reps >= 1");
  size_t bytes = (size_t)elements * 16;
  int nb = (int)((elements + 255) / 256);
  StreamPtrs a{};
  double2* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LCEQ_OK;
  for (int k = 0; k < streams && rc == LCEQ_OK; ++k) {
    void* q = nullptr;
    if (hipMalloc(&q, bytes) != hipSuccess || hipMemset(q, 0, bytes) != hipSuccess)
      rc = fail(LCEQ_ERR_HIP, "stream buffer %d could not be set up", k);
    a.p[k] = static_cast<const double2*>(q);
  }
  if (rc == LCEQ_OK && (hipMalloc(&out, bytes) != hipSuccess ||
                        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(LCEQ_ERR_HIP, "stream output or events could not be set up");
  if (rc == LCEQ_OK) {
    auto launch = [&]() {
      switch (streams) {
        case 8: hipLaunchKernelGGL(k_streams<8>, dim3(nb), dim3(256), 0, 0, a, out, elements); break;
        case 16: hipLaunchKernelGGL(k_streams<16>, dim3(nb), dim3(256), 0, 0, a, out, elements); break;
        default: hipLaunchKernelGGL(k_streams<30>, dim3(nb), dim3(256), 0, 0, a, out, elements); break;
      }
    };
    for (int r = 0; r < 3; ++r) launch();
    float el = 0.0f;
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess)
      rc = fail(LCEQ_ERR_HIP, "stream timing failed");
    else
      *us_per_pass = 1e3 * (double)el / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (int k = 0; k < streams; ++k)
    if (a.p[k]) (void)hipFree(const_cast<double2*>(a.p[k]));
  if (out) (void)hipFree(out);
  return rc;
}

// Time `trials` production sweeps (schedule 0, no measurement) at every
// launch shape and keep the fastest; the fields, observables, pass counter
// and history are restored, so tuning leaves the chain where it was.
extern "C" int lceq_autotune(lceq_lattice* L, int n_or, int trials, double* ms) {
  NEEDL(L);
  if (trials < 1 || n_or < 0) return fail(LCEQ_ERR_ARG, "bad autotune arguments");
  if (L->rng_mode != LCEQ_RNG_PHILOX) return fail(LCEQ_ERR_ARG, "autotune needs Philox streams");
  size_t np = (size_t)L->g.npad;
  std::vector<double> u(24 * np), a(6 * np), o(LCEQ_OBS_N);
  HIPCHK(hipMemcpyAsync(u.data(), L->U, u.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipMemcpyAsync(a.data(), L->A, a.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipMemcpyAsync(o.data(), L->obs, o.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  uint64_t pass0 = L->pass;
  int rows0 = L->hist_rows;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(LCEQ_ERR_HIP, "hipEventCreate failed");
  }
  const int shapes[4] = {64, 128, 256, 512};
  double best = 1e300, delta = o[OBS_DELTA];
  int best_tb = L->tb, best_x = L->xcd, rc = LCEQ_OK;
  for (int i = 0; i < 8 && rc == LCEQ_OK; ++i) {
    L->tb = shapes[i / 2];
    L->xcd = i % 2;
    if ((rc = lceq_sweep(L, LCEQ_SWEEP_KP_OR_METRO, 1, n_or, &delta, 0))) break;
    (void)hipEventRecord(e0, L->stream);
    if ((rc = lceq_sweep(L, LCEQ_SWEEP_KP_OR_METRO, trials, n_or, &delta, 0))) break;
    (void)hipEventRecord(e1, L->stream);
    float el = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess) {
      rc = fail(LCEQ_ERR_HIP, "timing failed");
      break;
    }
    double per = (double)el / trials;
    if (ms) ms[i] = per;
    if (per < best) {
      best = per;
      best_tb = L->tb;
      best_x = L->xcd;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  L->tb = best_tb;
  L->xcd = best_x;
  L->pass = pass0;
  L->hist_rows = rows0;
  HIPCHK(hipMemcpyAsync(L->U, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(L->A, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(L->obs, o.data(), o.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return rc;
}

extern "C" int lceq_set_site_order(lceq_lattice* L, int strip_rows) {
  NEEDL(L);
  int sy = strip_rows;
  if (sy != 0 && (sy < 2 || (sy & (sy - 1)) || L->g.ly % sy))
    return fail(LCEQ_ERR_ARG, "strip rows must be 0 or a power of two >= 2 dividing ly = %d",
                L->g.ly);
  // order_site divides a launch position by lx/2 to get its row; the
  // reciprocal division is exact for quotients below 2^22, so the rows
  // ly * lz of a parity must stay below that
  if (sy != 0 && (int64_t)L->g.ly * L->g.lz >= ((int64_t)1 << 22))
    return fail(LCEQ_ERR_ARG, "strip rows need ly * lz < 2^22 (ly = %d, lz = %d)", L->g.ly,
                L->g.lz);
  HIPCHK(hipStreamSynchronize(L->stream));
  L->strip = sy;
  return LCEQ_OK;
}

extern "C" int lceq_get_site_order(lceq_lattice* L, int* strip_rows) {
  NEEDL(L);
  if (strip_rows) *strip_rows = L->strip;
  return LCEQ_OK;
}

// Time `trials` production sweeps with each site order of 0, 4, 8 and 16
// strip rows (those dividing ly) at the current launch shape and keep the
// fastest; the fields, the observables, the pass counter and the history are
// restored.  ms (or NULL) receives 4 times per sweep, -1 for an order that
// does not fit the lattice.
extern "C" int lceq_autotune_site_order(lceq_lattice* L, int n_or, int trials, double* ms) {
  NEEDL(L);
  if (trials < 1 || n_or < 0) return fail(LCEQ_ERR_ARG, "bad autotune arguments");
  if (L->rng_mode != LCEQ_RNG_PHILOX) return fail(LCEQ_ERR_ARG, "autotune needs Philox streams");
  size_t np = (size_t)L->g.npad;
  std::vector<double> u(24 * np), a(6 * np), o(LCEQ_OBS_N);
  HIPCHK(hipMemcpyAsync(u.data(), L->U, u.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipMemcpyAsync(a.data(), L->A, a.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipMemcpyAsync(o.data(), L->obs, o.size() * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  uint64_t pass0 = L->pass;
  int rows0 = L->hist_rows;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(LCEQ_ERR_HIP, "hipEventCreate failed");
  }
  const int orders[4] = {0, 4, 8, 16};
  double best = 1e300, delta = o[OBS_DELTA];
  int best_sy = L->strip, rc = LCEQ_OK;
  for (int i = 0; i < 4 && rc == LCEQ_OK; ++i) {
    int sy = orders[i];
    if (sy && (L->g.ly % sy || (int64_t)L->g.ly * L->g.lz >= ((int64_t)1 << 22))) {
      if (ms) ms[i] = -1.0;  // the order does not fit (lceq_set_site_order's limits)
      continue;
    }
    L->strip = sy;
    if ((rc = lceq_sweep(L, LCEQ_SWEEP_KP_OR_METRO, 1, n_or, &delta, 0))) break;
    (void)hipEventRecord(e0, L->stream);
    if ((rc = lceq_sweep(L, LCEQ_SWEEP_KP_OR_METRO, trials, n_or, &delta, 0))) break;
    (void)hipEventRecord(e1, L->stream);
    float el = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess) {
      rc = fail(LCEQ_ERR_HIP, "timing failed");
      break;
    }
    double per = (double)el / trials;
    if (ms) ms[i] = per;
    if (per < best) {
      best = per;
      best_sy = sy;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  L->strip = best_sy;
  L->pass = pass0;
  L->hist_rows = rows0;
  HIPCHK(hipMemcpyAsync(L->U, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(L->A, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(L->obs, o.data(), o.size() * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return rc;
}

extern "C" int lceq_set_imode(lceq_lattice* L, int imode) {
  NEEDL(L);
  if (imode != 0 && imode != 1) return fail(LCEQ_ERR_ARG, "imode must be 0 or 1");
  L->imode = imode;
  return LCEQ_OK;
}

// -------------------------------------------------------------- uniforms

extern "C" int lceq_rng_philox(lceq_lattice* L, uint64_t seed, uint64_t pass_counter) {
  NEEDL(L);
  L->rng_mode = LCEQ_RNG_PHILOX;
  L->seed = seed;
  L->pass = pass_counter;
  return LCEQ_OK;
}

extern "C" int lceq_rng_feed(lceq_lattice* L, const double* uniforms, int64_t n) {
  NEEDL(L);
  if (n < 0 || (n > 0 && !uniforms)) return fail(LCEQ_ERR_ARG, "bad fed buffer");
  if (n > L->fed_cap) {
    if (L->fed) HIPCHK(hipFree(L->fed));
    L->fed = nullptr;
    L->fed_cap = 0;
    HIPCHK(hipMalloc((void**)&L->fed, (size_t)(n > 0 ? n : 1) * sizeof(double)));
    L->fed_cap = n;
  }
  if (n > 0)
    HIPCHK(hipMemcpyAsync(L->fed, uniforms, (size_t)n * sizeof(double), hipMemcpyHostToDevice,
                          L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  L->fed_n = n;
  L->fed_pos = 0;
  L->rng_mode = LCEQ_RNG_FED;
  return LCEQ_OK;
}

extern "C" int lceq_rng_state(lceq_lattice* L, int64_t* fed_consumed, uint64_t* pass_counter) {
  NEEDL(L);
  if (fed_consumed) *fed_consumed = L->fed_pos;
  if (pass_counter) *pass_counter = L->pass;
  return LCEQ_OK;
}

// Reserve `count` fed uniforms (fed mode) and return their device address,
// or null in philox mode.  Fails when the fed stream is exhausted.
static int take_uniforms(lceq_lattice* L, int64_t count, const double** ptr) {
  *ptr = nullptr;
  if (L->rng_mode != LCEQ_RNG_FED) return LCEQ_OK;
  if (L->fed_pos + count > L->fed_n)
    return fail(LCEQ_ERR_RNG, "fed uniform stream exhausted: need %lld, have %lld",
                (long long)count, (long long)(L->fed_n - L->fed_pos));
  *ptr = L->fed + L->fed_pos;
  L->fed_pos += count;
  return LCEQ_OK;
}

// next pass number for philox mode (fed mode leaves the counter alone)
uint64_t next_pass(lceq_lattice* L) {
  uint64_t p = L->pass;
  if (L->rng_mode == LCEQ_RNG_PHILOX) L->pass += 1;
  return p;
}

// ---------------------------------------------------------------- fields

static int ensure_scratch(lceq_lattice* L, size_t bytes) {
  if (bytes <= L->scratch_bytes) return LCEQ_OK;
  if (L->scratch) HIPCHK(hipFree(L->scratch));
  L->scratch = nullptr;
  L->scratch_bytes = 0;
  HIPCHK(hipMalloc((void**)&L->scratch, bytes));
  L->scratch_bytes = bytes;
  return LCEQ_OK;
}

// host reference layout <-> device layout (field_offset).  ncomp = 4 (links)
// or 3 (Higgs), nclass = 6 (links: mu*2+p) or 2 (Higgs: p).
static int put_field(lceq_lattice* L, double* dev, const double* host, int64_t ld, int ncomp,
                     int nclass) {
  NEEDL(L);
  const Geom& g = L->g;
  if (!host || ld < g.n) return fail(LCEQ_ERR_ARG, "bad host array or ld < N");
  std::vector<double> st((size_t)nclass * ncomp * g.npad, 0.0);
  for (int cls = 0; cls < nclass; ++cls)
    for (int s = 0; s < g.n; ++s)
      for (int c = 0; c < ncomp; ++c)
        st[field_offset(ncomp, cls, c, s, g.npad)] = host[((size_t)cls * ld + s) * ncomp + c];
  HIPCHK(hipMemcpyAsync(dev, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

static int get_field(lceq_lattice* L, const double* dev, double* host, int64_t ld, int ncomp,
                     int nclass) {
  NEEDL(L);
  const Geom& g = L->g;
  if (!host || ld < g.n) return fail(LCEQ_ERR_ARG, "bad host array or ld < N");
  std::vector<double> st((size_t)nclass * ncomp * g.npad);
  HIPCHK(hipMemcpyAsync(st.data(), dev, st.size() * sizeof(double), hipMemcpyDeviceToHost,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  for (int cls = 0; cls < nclass; ++cls)
    for (int s = 0; s < g.n; ++s)
      for (int c = 0; c < ncomp; ++c)
        host[((size_t)cls * ld + s) * ncomp + c] = st[field_offset(ncomp, cls, c, s, g.npad)];
  return LCEQ_OK;
}

extern "C" int lceq_put_links(lceq_lattice* L, const double* u, int64_t ld) {
  NEEDL(L);
  return put_field(L, L->U, u, ld, 4, 6);
}
extern "C" int lceq_get_links(lceq_lattice* L, double* u, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->U, u, ld, 4, 6);
}
extern "C" int lceq_put_higgs(lceq_lattice* L, const double* a, int64_t ld) {
  NEEDL(L);
  return put_field(L, L->A, a, ld, 3, 2);
}
extern "C" int lceq_get_higgs(lceq_lattice* L, double* a, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->A, a, ld, 3, 2);
}
extern "C" int lceq_put_links_fixed(lceq_lattice* L, const double* u, int64_t ld) {
  NEEDL(L);
  return put_field(L, L->Uf, u, ld, 4, 6);
}
extern "C" int lceq_get_links_fixed(lceq_lattice* L, double* u, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->Uf, u, ld, 4, 6);
}
extern "C" int lceq_put_higgs_fixed(lceq_lattice* L, const double* a, int64_t ld) {
  NEEDL(L);
  return put_field(L, L->Af, a, ld, 3, 2);
}
extern "C" int lceq_get_higgs_fixed(lceq_lattice* L, double* a, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->Af, a, ld, 3, 2);
}
extern "C" int lceq_get_staple(lceq_lattice* L, double* v, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->Vst, v, ld, 4, 6);
}
extern "C" int lceq_get_force(lceq_lattice* L, double* f, int64_t ld) {
  NEEDL(L);
  return get_field(L, L->F, f, ld, 3, 1);
}

extern "C" int lceq_unitconf(lceq_lattice* L) {
  NEEDL(L);
  launch_unitconf(L);
  return check_launch();
}
extern "C" int lceq_testconf(lceq_lattice* L) {
  NEEDL(L);
  launch_testconf(L);
  return check_launch();
}
extern "C" int lceq_hotstart(lceq_lattice* L, uint64_t seed, double sigma) {
  NEEDL(L);
  launch_hotstart(L, seed, sigma);
  return check_launch();
}

int read_obs(lceq_lattice* L, double* host) {
  HIPCHK(hipMemcpyAsync(host, L->obs, LCEQ_OBS_N * sizeof(double), hipMemcpyDeviceToHost,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

int write_obs(lceq_lattice* L, int index, double value) {
  HIPCHK(hipMemcpyAsync(L->obs + index, &value, sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

// ------------------------------------------------------------- hot path

extern "C" int lceq_staples(lceq_lattice* L, int eo, int mu) {
  NEEDL(L);
  if ((eo != 0 && eo != 1) || mu < 1 || mu > 3) return fail(LCEQ_ERR_ARG, "eo in {0,1}, mu in 1..3");
  launch_staples(L, eo, mu - 1);
  return check_launch();
}

extern "C" int lceq_heat_kpi(lceq_lattice* L, int n, const double* b, double* v,
                             const double* a0_old, int* nacch) {
  NEEDL(L);
  if (n < 0 || (n > 0 && (!b || !v || !a0_old))) return fail(LCEQ_ERR_ARG, "bad arrays");
  if (n == 0) {
    if (nacch) *nacch = 0;
    return LCEQ_OK;
  }
  int nb = blocks_for(n);
  if (nb > L->pstride * LCEQ_NSLOT * LCEQ_NQ) return fail(LCEQ_ERR_ARG, "n too large for lattice");
  int rc = ensure_scratch(L, (size_t)6 * n * sizeof(double));
  if (rc) return rc;
  const double* fed;
  if ((rc = take_uniforms(L, 6 * (int64_t)n, &fed))) return rc;
  double *db = L->scratch, *da0 = db + n, *dv = da0 + n;
  HIPCHK(hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(da0, a0_old, n * sizeof(double), hipMemcpyHostToDevice, L->stream));
  launch_heat_kpi(L, n, db, da0, dv, fed, next_pass(L));
  if ((rc = check_launch())) return rc;
  std::vector<double> part(nb);
  HIPCHK(hipMemcpyAsync(v, dv, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipMemcpyAsync(part.data(), L->partials, nb * sizeof(double), hipMemcpyDeviceToHost,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  double cnt = 0.0;
  for (double x : part) cnt += x;
  if (nacch) *nacch = (int)cnt;
  return LCEQ_OK;
}

// one full link sweep (eo outer, mu inner as in f/agauge.f:53-57); with
// reduce the six launches leave their plaquette and acceptance partial sums
// in slots 0-5 for launch_finalize
int link_sweep(lceq_lattice* L, int kp, int measflag, int renorm_on_write, int reduce) {
  int64_t per = (kp ? 7 : 1) * (int64_t)L->g.n;
  const double* fed0;
  int rc = take_uniforms(L, 6 * per, &fed0);
  if (rc) return rc;
  int kc = kp ? KC_LINK_KP : KC_LINK_OR;
  prof_begin(L, kc);
  for (int eo = 0; eo < 2; ++eo)
    for (int mu = 0; mu < 3; ++mu) {
      int slot = eo * 3 + mu;
      launch_link(L, kp, eo, mu, measflag, renorm_on_write, fed0 ? fed0 + slot * per : nullptr,
                  next_pass(L), reduce);
    }
  prof_end(L, kc);
  return check_launch();
}

static int link_call(lceq_lattice* L, int kp, int measflag, double* accept) {
  NEEDL(L);
  int rc = link_sweep(L, kp, measflag, 0, 1);
  if (rc) return rc;
  launch_finalize(L, kp ? 1 : 2, measflag, 0, 0, nullptr);
  if (accept) {
    double o[LCEQ_OBS_N];
    if ((rc = read_obs(L, o))) return rc;
    *accept = o[OBS_ACC_LINK];
  }
  return LCEQ_OK;
}

extern "C" int lceq_agauge(lceq_lattice* L, int measflag, double* accept) {
  return link_call(L, 1, measflag, accept);
}

extern "C" int lceq_link_or(lceq_lattice* L, int measflag, double* accept) {
  return link_call(L, 0, measflag, accept);
}

extern "C" int lceq_a_force(lceq_lattice* L, int eo) {
  NEEDL(L);
  if (eo != 0 && eo != 1) return fail(LCEQ_ERR_ARG, "eo in {0,1}");
  launch_force(L, eo);
  return check_launch();
}

extern "C" int lceq_gauss(lceq_lattice* L, double* x, const double* r, int nh) {
  NEEDL(L);
  if (nh < 0 || (nh > 0 && (!x || !r))) return fail(LCEQ_ERR_ARG, "bad arrays");
  if (nh == 0) return LCEQ_OK;
  int rc = ensure_scratch(L, (size_t)2 * nh * sizeof(double));
  if (rc) return rc;
  const double* fed;
  if ((rc = take_uniforms(L, 2 * (int64_t)nh, &fed))) return rc;
  HIPCHK(hipMemcpyAsync(L->scratch, r, nh * sizeof(double), hipMemcpyHostToDevice, L->stream));
  launch_gauss(L, nh, L->scratch, L->scratch + nh, fed, next_pass(L));
  if ((rc = check_launch())) return rc;
  HIPCHK(hipMemcpyAsync(x, L->scratch + nh, nh * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

// both parities of one Higgs routine (kind 1/2/3 = ahiggs1/2/3); partial
// sums land in slots 6-7 for launch_finalize
static int higgs_sweep(lceq_lattice* L, int kind, int measflag) {
  int64_t per = (kind == 1 ? 4 : kind == 2 ? 7 : 1) * (int64_t)L->g.n;
  const double* fed0;
  int rc = take_uniforms(L, 2 * per, &fed0);
  if (rc) return rc;
  prof_begin(L, KC_HIGGS);
  for (int eo = 0; eo < 2; ++eo)
    launch_higgs(L, kind, eo, measflag, fed0 ? fed0 + eo * per : nullptr, next_pass(L));
  prof_end(L, KC_HIGGS);
  return check_launch();
}

extern "C" int lceq_ahiggs1(lceq_lattice* L, double* delta, int measflag, double* accept) {
  NEEDL(L);
  if (!delta) return fail(LCEQ_ERR_ARG, "delta is required");
  int rc = write_obs(L, OBS_DELTA, *delta);
  if (rc) return rc;
  if ((rc = higgs_sweep(L, 1, measflag))) return rc;
  launch_finalize(L, 0, 0, 1, measflag, nullptr);
  double o[LCEQ_OBS_N];
  if ((rc = read_obs(L, o))) return rc;
  *delta = o[OBS_DELTA];
  if (accept) *accept = o[OBS_ACC_HIGGS];
  return LCEQ_OK;
}

static int higgs_call(lceq_lattice* L, int kind, int measflag, double* accept) {
  NEEDL(L);
  int rc = higgs_sweep(L, kind, measflag);
  if (rc) return rc;
  launch_finalize(L, 0, 0, kind, measflag, nullptr);
  if (accept && measflag == 1) {
    double o[LCEQ_OBS_N];
    if ((rc = read_obs(L, o))) return rc;
    *accept = o[OBS_ACC_HIGGS];
  }
  return LCEQ_OK;
}
extern "C" int lceq_ahiggs2(lceq_lattice* L, int measflag, double* accept) {
  return higgs_call(L, 2, measflag, accept);
}
extern "C" int lceq_ahiggs3(lceq_lattice* L, int measflag, double* accept) {
  return higgs_call(L, 3, measflag, accept);
}

extern "C" int lceq_renorm(lceq_lattice* L) {
  NEEDL(L);
  launch_renorm(L, L->U);
  return check_launch();
}
extern "C" int lceq_check_fields(lceq_lattice* L, double tol, double* max_dev,
                                 int64_t* nonfinite) {
  NEEDL(L);
  int nb = blocks_for(L->g.n);
  int rc = ensure_scratch(L, (size_t)2 * nb * sizeof(double));
  if (rc) return rc;
  launch_check_fields(L, L->scratch);
  if ((rc = check_launch())) return rc;
  std::vector<double> h(2 * (size_t)nb);
  HIPCHK(hipMemcpyAsync(h.data(), L->scratch, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  double dev = 0.0, bad = 0.0;
  for (int b = 0; b < nb; ++b) {
    dev = std::max(dev, h[2 * b]);
    bad += h[2 * b + 1];
  }
  if (max_dev) *max_dev = dev;
  if (nonfinite) *nonfinite = (int64_t)bad;
  if (bad > 0.0) return fail(LCEQ_ERR_FIELD, "%.0f non-finite link or Higgs values", bad);
  if (tol >= 0.0 && dev > tol)
    return fail(LCEQ_ERR_FIELD, "links off SU(2) by %.3g (tolerance %.3g)", dev, tol);
  return LCEQ_OK;
}

extern "C" int lceq_u_f_renorm(lceq_lattice* L) {
  NEEDL(L);
  launch_renorm(L, L->Uf);
  return check_launch();
}

extern "C" int lceq_or_gauge_fixing(lceq_lattice* L, double* da1) {
  NEEDL(L);
  launch_gfix(L, 0, 0, 0, 0);
  launch_gfix(L, 1, 1, 0, 0);
  launch_gfix_finalize(L, 0);
  int rc = check_launch();
  if (rc) return rc;
  if (da1) {
    double o[LCEQ_OBS_N];
    if ((rc = read_obs(L, o))) return rc;
    *da1 = o[OBS_DA];
  }
  return LCEQ_OK;
}

extern "C" int lceq_landau_gauge(lceq_lattice* L, double* da1, int* iter_or1) {
  NEEDL(L);
  size_t np = (size_t)L->g.npad;
  HIPCHK(hipMemcpyAsync(L->Uf, L->U, 24 * np * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(L->Af, L->A, 6 * np * sizeof(double), hipMemcpyDeviceToDevice, L->stream));
  double init[3] = {10000.0, 0.0, 0.0};  // da1 = 10000 (f/landau_gauge.f:28), iter, done
  HIPCHK(hipMemcpyAsync(L->obs + OBS_DA, init, 2 * sizeof(double), hipMemcpyHostToDevice, L->stream));
  double z2[2] = {0.0, 0.0};
  HIPCHK(hipMemcpyAsync(L->obs + OBS_GF_DONE, z2, 2 * sizeof(double), hipMemcpyHostToDevice,
                        L->stream));
  // Batches of guarded iterations; after convergence (or the cap) the
  // remaining launches of a batch return at once.  Inside a batch the loop
  // condition for the iterations so far is tested by the next iteration's
  // first launch; one finalize tests it after the batch's last iteration.
  // With more than 2048 blocks per launch every block's re-sum of all the
  // partials costs more than a one-block test launch per iteration (26 us
  // of a 352 us iteration at 160^3), so there each iteration is followed by
  // the finalize instead: the same sum in the same order, the same stop.
  const bool test_launch = blocks_for(L->g.n) > 2048;
  int batch = 8, done_iter = 0;
  double o[LCEQ_OBS_N];
  for (;;) {
    for (int k = 0; k < batch; ++k) {
      launch_gfix(L, 0, 0, 0, 1, test_launch ? 0 : k > 0, done_iter);
      launch_gfix(L, 1, 1, 1, 1);
      ++done_iter;
      if (test_launch) launch_gfix_finalize(L, 1, done_iter);
    }
    if (!test_launch) launch_gfix_finalize(L, 1, done_iter);
    int rc = check_launch();
    if (rc) return rc;
    if ((rc = read_obs(L, o))) return rc;
    if (o[OBS_GF_DONE] != 0.0) break;
    if (batch < 256) batch *= 2;
  }
  if (da1) *da1 = o[OBS_DA];
  if (iter_or1) *iter_or1 = (int)o[OBS_ITER];
  if (o[OBS_GF_CAPPED] != 0.0)
    return fail(LCEQ_ERR_NOT_CONVERGED, "landau_gauge: da = %g after max_iter = %d iterations",
                o[OBS_DA], L->max_iter);
  return LCEQ_OK;
}

static int run_corr(lceq_lattice* L, std::vector<double>& c) {
  launch_slices(L);
  launch_corr(L);
  int rc = check_launch();
  if (rc) return rc;
  c.resize(5 * (size_t)L->g.lz + 5);
  HIPCHK(hipMemcpyAsync(c.data(), L->corr, c.size() * sizeof(double), hipMemcpyDeviceToHost,
                        L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

extern "C" int lceq_correlator(lceq_lattice* L, double* corr_a, double* corr_u, double* corr_3) {
  NEEDL(L);
  std::vector<double> c;
  int rc = run_corr(L, c);
  if (rc) return rc;
  int lz = L->g.lz;
  if (corr_a) memcpy(corr_a, c.data(), lz * sizeof(double));
  if (corr_u) memcpy(corr_u, c.data() + lz, lz * sizeof(double));
  if (corr_3) memcpy(corr_3, c.data() + 2 * lz, lz * sizeof(double));
  HIPCHK(hipMemcpyAsync(L->obs + OBS_A0F, c.data() + 5 * lz + 2, 3 * sizeof(double),
                        hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

extern "C" int lceq_scalarcorr(lceq_lattice* L, double* mean, double* corr) {
  NEEDL(L);
  std::vector<double> c;
  int rc = run_corr(L, c);
  if (rc) return rc;
  int lz = L->g.lz;
  if (mean) *mean = c[5 * lz];
  if (corr) memcpy(corr, c.data() + 3 * lz, lz * sizeof(double));
  return LCEQ_OK;
}

extern "C" int lceq_polyakov(lceq_lattice* L, double* mean_sq, double* corr) {
  NEEDL(L);
  std::vector<double> c;
  int rc = run_corr(L, c);
  if (rc) return rc;
  int lz = L->g.lz;
  if (mean_sq) *mean_sq = c[5 * lz + 1];
  if (corr) memcpy(corr, c.data() + 4 * lz, lz * sizeof(double));
  return LCEQ_OK;
}

extern "C" int lceq_multall(lceq_lattice* L, int kind, const double* op1, const double* op2,
                            double* out, int n, const int* s1, const int* s2) {
  NEEDL(L);
  if (kind < 0 || kind > 5 || n < 0 || (n > 0 && (!op1 || !op2 || !out)))
    return fail(LCEQ_ERR_ARG, "bad multall arguments");
  if (n == 0) return LCEQ_OK;
  // operand extents: the largest site index referenced
  int64_t m1 = n, m2 = n;
  for (int i = 0; i < n; ++i) {
    if (s1 && (s1[i] < 0)) return fail(LCEQ_ERR_ARG, "negative site index");
    if (s2 && (s2[i] < 0)) return fail(LCEQ_ERR_ARG, "negative site index");
    if (s1 && (int64_t)s1[i] + 1 > m1) m1 = (int64_t)s1[i] + 1;
    if (s2 && (int64_t)s2[i] + 1 > m2) m2 = (int64_t)s2[i] + 1;
  }
  int w1 = kind >= 4 ? 3 : 4;
  size_t bytes = ((size_t)w1 * m1 + 4 * (size_t)m2 + 4 * (size_t)n) * sizeof(double) +
                 2 * (size_t)n * sizeof(int) + 64;
  int rc = ensure_scratch(L, bytes);
  if (rc) return rc;
  double* d1 = L->scratch;
  double* d2 = d1 + (size_t)w1 * m1;
  double* dout = d2 + 4 * (size_t)m2;
  int* ds1 = (int*)(dout + 4 * (size_t)n);
  int* ds2 = ds1 + n;
  HIPCHK(hipMemcpyAsync(d1, op1, (size_t)w1 * m1 * sizeof(double), hipMemcpyHostToDevice, L->stream));
  HIPCHK(hipMemcpyAsync(d2, op2, 4 * (size_t)m2 * sizeof(double), hipMemcpyHostToDevice, L->stream));
  if (s1) HIPCHK(hipMemcpyAsync(ds1, s1, n * sizeof(int), hipMemcpyHostToDevice, L->stream));
  if (s2) HIPCHK(hipMemcpyAsync(ds2, s2, n * sizeof(int), hipMemcpyHostToDevice, L->stream));
  launch_multall(L, kind, d1, d2, dout, n, s1 ? ds1 : nullptr, s2 ? ds2 : nullptr);
  if ((rc = check_launch())) return rc;
  HIPCHK(hipMemcpyAsync(out, dout, 4 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, L->stream));
  HIPCHK(hipStreamSynchronize(L->stream));
  return LCEQ_OK;
}

extern "C" int lceq_observables(lceq_lattice* L, double* out) {
  NEEDL(L);
  if (!out) return fail(LCEQ_ERR_ARG, "null output");
  double o[LCEQ_OBS_N];
  int rc = read_obs(L, o);
  if (rc) return rc;
  for (int i = 0; i < 7; ++i) out[i] = o[OBS_PLAQ0 + i];
  return LCEQ_OK;
}

// ----------------------------------------------------------------- sweeps

extern "C" int lceq_sweep(lceq_lattice* L, int schedule, int nsweep, int n_or, double* delta,
                          int measflag) {
  NEEDL(L);
  if (nsweep < 0 || n_or < 0) return fail(LCEQ_ERR_ARG, "nsweep, n_or must be >= 0");
  if (schedule != LCEQ_SWEEP_KP_OR_METRO && schedule != LCEQ_SWEEP_REFERENCE)
    return fail(LCEQ_ERR_ARG, "unknown schedule %d", schedule);
  int rc;
  if (schedule == LCEQ_SWEEP_KP_OR_METRO) {
    if (!delta) return fail(LCEQ_ERR_ARG, "delta is required for schedule 0");
    if ((rc = write_obs(L, OBS_DELTA, *delta))) return rc;
  }
  if (measflag && L->hist_rows + nsweep > L->hist_cap)
    return fail(LCEQ_ERR_ARG, "history full (%d rows); call lceq_history", L->hist_cap);
  for (int it = 0; it < nsweep; ++it) {
    double* row = measflag ? L->hist + (size_t)(L->hist_rows++) * LCEQ_OBS_N : nullptr;
    if (schedule == LCEQ_SWEEP_KP_OR_METRO) {
      // only the last link sweep carries the measurement; the ahiggs1 step
      // adapts every sweep, so the Higgs launches always reduce.  One
      // finalize launch closes the sweep.
      int meas_kp = measflag && n_or == 0;
      if ((rc = link_sweep(L, 1, meas_kp, 1, meas_kp))) return rc;
      prof_begin(L, KC_LINK_OR);
      for (int k = 0; k < n_or; ++k) {
        int m = measflag && k == n_or - 1;
        if ((rc = link_sweep(L, 0, m, 1, m))) break;
      }
      prof_end(L, KC_LINK_OR);
      if (rc) return rc;
      if ((rc = higgs_sweep(L, 1, measflag))) return rc;
      launch_finalize(L, measflag ? (n_or ? 2 : 1) : 0, measflag, 1, measflag, row);
    } else {  // f/main.f:38-41
      if ((rc = higgs_sweep(L, 2, measflag))) return rc;
      if (measflag) launch_finalize(L, 0, 0, 2, 1, nullptr);
      if ((rc = higgs_sweep(L, 3, measflag))) return rc;
      if ((rc = link_sweep(L, 1, measflag, 0, measflag))) return rc;
      if (measflag) launch_finalize(L, 1, 1, 3, 1, row);
      launch_renorm(L, L->U);
    }
  }
  if ((rc = check_launch())) return rc;
  if (schedule == LCEQ_SWEEP_KP_OR_METRO) {
    HIPCHK(hipMemcpyAsync(delta, L->obs + OBS_DELTA, sizeof(double), hipMemcpyDeviceToHost,
                          L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
  }
  return LCEQ_OK;
}

extern "C" int lceq_history(lceq_lattice* L, double* out, int n) {
  NEEDL(L);
  if (n < 0 || n > L->hist_rows) return fail(LCEQ_ERR_ARG, "only %d rows recorded", L->hist_rows);
  if (n > 0) {
    if (!out) return fail(LCEQ_ERR_ARG, "null output");
    HIPCHK(hipMemcpyAsync(out, L->hist + (size_t)(L->hist_rows - n) * LCEQ_OBS_N,
                          (size_t)n * LCEQ_OBS_N * sizeof(double), hipMemcpyDeviceToHost,
                          L->stream));
    HIPCHK(hipStreamSynchronize(L->stream));
  }
  L->hist_rows = 0;
  return LCEQ_OK;
}

// ---------------------------------------------------------------- timing

extern "C" int lceq_profile_begin(lceq_lattice* L, int kernel_class, int max_launches) {
  NEEDL(L);
  if (kernel_class < 1 || kernel_class > 4 || max_launches < 1)
    return fail(LCEQ_ERR_ARG, "bad profile arguments");
  int cap = 2 * max_launches;
  if (cap > L->prof_cap) {
    if (L->prof_ev) {
      for (int i = 0; i < L->prof_cap; ++i) (void)hipEventDestroy(L->prof_ev[i]);
      delete[] L->prof_ev;
    }
    L->prof_ev = new hipEvent_t[cap];
    L->prof_cap = 0;
    for (int i = 0; i < cap; ++i) {
      HIPCHK(hipEventCreate(&L->prof_ev[i]));
      L->prof_cap = i + 1;
    }
  }
  L->prof_n = 0;
  L->prof_depth = 0;
  L->prof_launches = 0;
  L->prof_class = kernel_class;
  return LCEQ_OK;
}

extern "C" int lceq_profile_end(lceq_lattice* L, double* total_ms, int* launches) {
  NEEDL(L);
  HIPCHK(hipStreamSynchronize(L->stream));
  double tot = 0.0;
  int pairs = L->prof_n / 2;
  for (int i = 0; i < pairs; ++i) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, L->prof_ev[2 * i], L->prof_ev[2 * i + 1]));
    tot += ms;
  }
  L->prof_class = 0;
  L->prof_n = 0;
  L->prof_depth = 0;
  if (total_ms) *total_ms = tot;
  if (launches) *launches = L->prof_launches;
  return LCEQ_OK;
}
