// lceq_mlp.hip -- learned proposal for the adjoint-Higgs update (no
// counterpart in the reference; the plain update it competes with is the
// Metropolis step of f/ahiggs1.f).
//
// Per site x of the parity being updated, with F = the hopping force of
// f/a_force.f (fixed while this parity is updated), a small network
//   10 inputs -> 64 SiLU -> 64 SiLU -> 6 outputs
// gives the mean mu and log-scale s of a Gaussian proposal in the frame
// (e1, e2, e3) with e1 = F/|F|.  The inputs are the local a (3) and the
// hopping sum (3) in that frame, |F|, beta_A, beta_2 and beta_4; the frame
// removes the global SO(3) orientation, on which the action does not depend.
// a' is drawn from N(mu(a), exp(s(a))) and accepted with the FP64 Metropolis-
// Hastings probability
//   min(1, exp(S(a) - S(a')) q(a | a') / q(a' | a)),
// S the site action of f/ahiggs1.f:64-92, so the chain is exact whatever the
// network computes; the reverse density needs a second evaluation at a'.
// Inference runs on bf16 MFMA (v_mfma_f32_16x16x32_bf16, fp32 accumulation),
// one 64-site batch per wave.  Training (maximum likelihood of the moves of
// several Metropolis steps at fixed F) is hand-written FP32 with Adam.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"
#include "lceq_sites.hpp"

using namespace lceq;

#include "lceq_mlp.hpp"

using namespace lceq::mlp;

namespace {

constexpr int TB = 256;

template <int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_higgs_mlp(MlpArgs p) {
  __shared__ MlpLds lds[4];
  mlp_higgs_site<ST>(p, xcd_block(p.nb), lds[threadIdx.x >> 6]);
}

// The inference network on given raw inputs (rows of NIN), for tests
__global__ __launch_bounds__(256) void k_mlp_eval(const MlpPack* __restrict__ P,
                                                  const float* __restrict__ x, float* __restrict__ y,
                                                  int n) {
  __shared__ MlpLds lds[4];
  int i = blockIdx.x * TB + threadIdx.x;
  int il = i < n ? i : n - 1;
  float xin[NIN], o[NOUT];
  MlpRegs W;
  mlp_load(P, W);
  mlp_stage_bias(P, lds[threadIdx.x >> 6]);
#pragma unroll
  for (int k = 0; k < NIN; ++k) xin[k] = x[(size_t)il * NIN + k];
  mlp_eval<4>(P, W, xin, o, lds[threadIdx.x >> 6]);
  if (i < n)
#pragma unroll
    for (int j = 0; j < NOUT; ++j) y[(size_t)i * NOUT + j] = o[j];
}

// ---------------------------------------------------------- training data

// Record of site s (every stride-th site of parity eo): the raw features of
// the current a at the current F, and (second kernel, after the heatbath
// steps) the new a in the same frame.
__global__ __launch_bounds__(256) void k_mlp_record(Geom g, const double* __restrict__ U,
                                                    const double* __restrict__ A, int eo, int stride,
                                                    int nrec, double ba, double b2, double b4,
                                                    float* __restrict__ out, int target) {
  int r = blockIdx.x * TB + threadIdx.x;
  if (r >= nrec) return;
  unsigned s = (unsigned)r * (unsigned)stride;
  Coord c = site_coord(g, eo, (int)s);
  vec3 F = hop_force(U, A, g, eo, s, c);
  Frame fr = make_frame(F.x);
  vec3 av = ld_higgs(A, g, eo, s);
  double y[3];
  to_frame(fr, av.x, y);
  float* rec = out + (size_t)r * NREC;
  if (target) {
    rec[NIN] = (float)y[0];
    rec[NIN + 1] = (float)y[1];
    rec[NIN + 2] = (float)y[2];
  } else {
    float x[NIN];
    raw_features(y, fr.fn, ba, b2, b4, x);
    for (int k = 0; k < NIN; ++k) rec[k] = x[k];
  }
}

// --------------------------------------------------------------- training

constexpr int TR_TB = 32;  // samples per block, one per lane

// Activations of the block's samples, [unit][sample] so that each lane walks
// its own column and the parameter-gradient sums walk contiguous rows
struct TrainLds {
  float x[NIN][TR_TB];
  float h1[NH][TR_TB], d1[NH][TR_TB], h2[NH][TR_TB], d2[NH][TR_TB];
  float dout[NOUT][TR_TB];
  float loss[TR_TB];
};

__device__ __forceinline__ float dsilu(float z) {
  float sg = 1.0f / (1.0f + expf(-z));
  return sg * (1.0f + z * (1.0f - sg));
}

// Gradient of the mean negative log-likelihood of one block of samples
// (records idx[b]) with respect to all parameters; one partial gradient and
// one partial loss per block.  Weights are read from global memory: every
// lane reads the same address at the same time.
__global__ __launch_bounds__(32) void k_mlp_grad(const float* __restrict__ w,
                                                 const float* __restrict__ norm,
                                                 const float* __restrict__ data,
                                                 const int* __restrict__ idx, int batch,
                                                 float* __restrict__ gpart,
                                                 float* __restrict__ lpart) {
  __shared__ TrainLds T;
  int t = threadIdx.x;
  int b = blockIdx.x * TR_TB + t;
  bool on = b < batch;
  const float* rec = data + (size_t)idx[on ? b : 0] * NREC;
  for (int k = 0; k < NIN; ++k) T.x[k][t] = (rec[k] - norm[k]) * norm[NIN + k];
  // forward; d1, d2 hold silu'(z) until the backward pass scales them
  for (int j = 0; j < NH; ++j) {
    float a = w[OFF_B1 + j];
    for (int k = 0; k < NIN; ++k) a += T.x[k][t] * w[OFF_W1 + k * NH + j];
    T.h1[j][t] = silu(a);
    T.d1[j][t] = dsilu(a);
  }
  for (int j = 0; j < NH; ++j) {
    float a = w[OFF_B2 + j];
    for (int k = 0; k < NH; ++k) a += T.h1[k][t] * w[OFF_W2 + k * NH + j];
    T.h2[j][t] = silu(a);
    T.d2[j][t] = dsilu(a);
  }
  float o[NOUT];
  for (int j = 0; j < NOUT; ++j) {
    float a = w[OFF_B3 + j];
    for (int k = 0; k < NH; ++k) a += T.h2[k][t] * w[OFF_W3 + k * NOUT + j];
    o[j] = a;
  }
  // loss sum_i [s_i + (y_i - mu_i)^2 exp(-2 s_i) / 2] and its gradient
  float scale = on ? 1.0f / (float)batch : 0.0f, loss = 0.0f;
  for (int i = 0; i < 3; ++i) {
    float sv = fminf(fmaxf(o[3 + i], S_MIN), S_MAX);
    float inv = expf(-sv), r = (rec[NIN + i] - o[i]) * inv;
    loss += sv + 0.5f * r * r;
    T.dout[i][t] = -r * inv * scale;
    bool clamped = o[3 + i] < S_MIN || o[3 + i] > S_MAX;
    T.dout[3 + i][t] = clamped ? 0.0f : (1.0f - r * r) * scale;
  }
  T.loss[t] = on ? loss : 0.0f;
  // backward through the hidden layers
  for (int k = 0; k < NH; ++k) {
    float a = 0.0f;
    for (int j = 0; j < NOUT; ++j) a += w[OFF_W3 + k * NOUT + j] * T.dout[j][t];
    T.d2[k][t] *= a;
  }
  for (int k = 0; k < NH; ++k) {
    float a = 0.0f;
    for (int j = 0; j < NH; ++j) a += w[OFF_W2 + k * NH + j] * T.d2[j][t];
    T.d1[k][t] *= a;
  }
  __syncthreads();
  // parameter gradients: sums over the block's samples in fixed order
  float* gp = gpart + (size_t)blockIdx.x * NPARAM;
  for (int i = t; i < NPARAM; i += TR_TB) {
    const float *u, *v;
    if (i < OFF_B1) {
      u = T.x[i / NH];
      v = T.d1[i % NH];
    } else if (i < OFF_W2) {
      u = nullptr;
      v = T.d1[i - OFF_B1];
    } else if (i < OFF_B2) {
      u = T.h1[(i - OFF_W2) / NH];
      v = T.d2[(i - OFF_W2) % NH];
    } else if (i < OFF_W3) {
      u = nullptr;
      v = T.d2[i - OFF_B2];
    } else if (i < OFF_B3) {
      u = T.h2[(i - OFF_W3) / NOUT];
      v = T.dout[(i - OFF_W3) % NOUT];
    } else {
      u = nullptr;
      v = T.dout[i - OFF_B3];
    }
    float a = 0.0f;
    if (u)
      for (int q = 0; q < TR_TB; ++q) a += u[q] * v[q];
    else
      for (int q = 0; q < TR_TB; ++q) a += v[q];
    gp[i] = a;
  }
  if (t == 0) {
    float l = 0.0f;
    for (int q = 0; q < TR_TB; ++q) l += T.loss[q];
    lpart[blockIdx.x] = l;
  }
}

// Sum the block gradients (fixed order) and take one Adam step
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ params, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ gpart, int nblk,
                                                  float lr, float c1, float c2) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= NPARAM) return;
  float g = 0.0f;
  for (int b = 0; b < nblk; ++b) g += gpart[(size_t)b * NPARAM + i];
  float mi = 0.9f * m[i] + 0.1f * g, vi = 0.999f * v[i] + 0.001f * g * g;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

uint64_t splitmix(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

unsigned short host_bf16(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

}  // namespace

struct lceq_mlp {
  int device;
  hipStream_t stream;
  std::vector<float> host;  // NSTATE: params, feature mean, inverse scale
  float *params, *adam_m, *adam_v, *norm, *gpart, *lpart;
  int gpart_blocks;
  int* idx;
  int idx_cap;
  MlpPack* pack;
  int sites_per_wave;  // learned Higgs launch: 0 = by lattice size, 32 or 64
  float* data;  // NREC floats per record
  int64_t nrec, cap;
  int64_t adam_t;
};

#define mfail lceq::set_error
#define MHIP(call)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess) return mfail(LCEQ_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

static int mlp_upload(lceq_mlp* M) {
  MHIP(hipMemcpyAsync(M->params, M->host.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice,
                      M->stream));
  MHIP(hipMemcpyAsync(M->norm, M->host.data() + NPARAM, 2 * NIN * sizeof(float),
                      hipMemcpyHostToDevice, M->stream));
  // inference pack
  MlpPack h;
  memset(&h, 0, sizeof(h));
  const float* w = M->host.data();
  // A operand of lane l: row (output unit) 16 nt + l % 16, K elements
  // 8 (l / 16) + j, natural for the inputs, hidden_of_k for hidden layers
  auto frag = [&](bf16x8& f, int K, int N, int off, int kb, int nt, int l, bool hidden) {
    unsigned short e[8];
    for (int j = 0; j < 8; ++j) {
      int g = l / 16, n = 16 * nt + l % 16;
      int k = hidden ? hidden_of_k(kb, g, j) : 32 * kb + 8 * g + j;
      e[j] = (k < K && n < N) ? host_bf16(w[off + k * N + n]) : (unsigned short)0;
    }
    memcpy(&f, e, 16);
  };
  for (int l = 0; l < 64; ++l) {
    for (int nt = 0; nt < 4; ++nt) {
      frag(h.w1[nt][l], NIN, NH, OFF_W1, 0, nt, l, false);
      frag(h.w2[nt][0][l], NH, NH, OFF_W2, 0, nt, l, true);
      frag(h.w2[nt][1][l], NH, NH, OFF_W2, 1, nt, l, true);
    }
    frag(h.w3[0][l], NH, NOUT, OFF_W3, 0, 0, l, true);
    frag(h.w3[1][l], NH, NOUT, OFF_W3, 1, 0, l, true);
  }
  for (int j = 0; j < NH; ++j) {
    h.b1[j] = w[OFF_B1 + j];
    h.b2[j] = w[OFF_B2 + j];
  }
  for (int j = 0; j < NOUT; ++j) h.b3[j] = w[OFF_B3 + j];
  for (int k = 0; k < NIN; ++k) {
    h.xmean[k] = w[NPARAM + k];
    h.xinv[k] = w[NPARAM + NIN + k];
  }
  MHIP(hipMemcpyAsync(M->pack, &h, sizeof(h), hipMemcpyHostToDevice, M->stream));
  MHIP(hipStreamSynchronize(M->stream));
  return LCEQ_OK;
}

extern "C" int lceq_mlp_create(uint64_t seed, lceq_mlp** out) {
  if (!out) return mfail(LCEQ_ERR_ARG, "null output");
  *out = nullptr;
  lceq_mlp* M = new lceq_mlp();
  M->host.assign(NSTATE, 0.0f);
  // Glorot-uniform weights, zero biases; a small last layer so that the
  // first proposals stay near the current configuration's scale
  uint64_t st = seed;
  auto uni = [&]() { return (double)(splitmix(st) >> 11) * 0x1.0p-53; };
  auto init = [&](int off, int fan_in, int fan_out, double gain) {
    double lim = gain * std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < fan_in * fan_out; ++i) M->host[off + i] = (float)((2.0 * uni() - 1.0) * lim);
  };
  init(OFF_W1, NIN, NH, 1.0);
  init(OFF_W2, NH, NH, 1.0);
  init(OFF_W3, NH, NOUT, 0.1);
  for (int k = 0; k < NIN; ++k) M->host[NPARAM + NIN + k] = 1.0f;  // identity normalisation
  int rc = LCEQ_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc) return;
    if (hipMalloc(p, bytes) != hipSuccess) rc = mfail(LCEQ_ERR_HIP, "hipMalloc(%zu) failed", bytes);
  };
  if (hipGetDevice(&M->device) != hipSuccess) rc = mfail(LCEQ_ERR_HIP, "no HIP device");
  if (!rc && hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess)
    rc = mfail(LCEQ_ERR_HIP, "hipStreamCreate failed");
  alloc((void**)&M->params, NPARAM * sizeof(float));
  alloc((void**)&M->adam_m, NPARAM * sizeof(float));
  alloc((void**)&M->adam_v, NPARAM * sizeof(float));
  alloc((void**)&M->norm, 2 * NIN * sizeof(float));
  alloc((void**)&M->pack, sizeof(MlpPack));
  if (rc) {
    lceq_mlp_destroy(M);
    return rc;
  }
  if ((rc = mlp_upload(M))) {
    lceq_mlp_destroy(M);
    return rc;
  }
  *out = M;
  return LCEQ_OK;
}

extern "C" int lceq_mlp_destroy(lceq_mlp* M) {
  if (!M) return LCEQ_OK;
  void* bufs[] = {M->params, M->adam_m, M->adam_v, M->norm, M->gpart, M->lpart, M->idx, M->pack,
                  M->data};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (M->stream) (void)hipStreamDestroy(M->stream);
  delete M;
  return LCEQ_OK;
}

extern "C" int lceq_mlp_get_state(lceq_mlp* M, float* state, int n) {
  if (!M || !state || n != NSTATE) return mfail(LCEQ_ERR_ARG, "state has %d floats", NSTATE);
  memcpy(state, M->host.data(), NSTATE * sizeof(float));
  return LCEQ_OK;
}

extern "C" int lceq_mlp_set_state(lceq_mlp* M, const float* state, int n) {
  if (!M || !state || n != NSTATE) return mfail(LCEQ_ERR_ARG, "state has %d floats", NSTATE);
  memcpy(M->host.data(), state, NSTATE * sizeof(float));
  M->adam_t = 0;
  MHIP(hipMemsetAsync(M->adam_m, 0, NPARAM * sizeof(float), M->stream));
  MHIP(hipMemsetAsync(M->adam_v, 0, NPARAM * sizeof(float), M->stream));
  return mlp_upload(M);
}

extern "C" int lceq_mlp_clear(lceq_mlp* M) {
  if (!M) return mfail(LCEQ_ERR_ARG, "null model");
  M->nrec = 0;
  return LCEQ_OK;
}

extern "C" int lceq_mlp_records(lceq_mlp* M, int64_t* n) {
  if (!M || !n) return mfail(LCEQ_ERR_ARG, "null argument");
  *n = M->nrec;
  return LCEQ_OK;
}

// Training pairs from the lattice's current configuration: for each parity,
// the features at the current F, then nsteps ahiggs1 Metropolis steps of
// that parity at the lattice's current step size (F does not change while
// one parity is updated), then the new a in the same frame; every stride-th
// site.  The Metropolis steps are kept (each is a valid update of the
// chain).  The Metropolis step rather than the ahiggs2 heatbath: the
// heatbath's Gaussian width beta_2 is negative in the broken phase.
const MlpPack* lceq::mlp::device_pack(const lceq_mlp* M) { return M->pack; }

float* lceq::mlp::reserve_records(lceq_mlp* M, int64_t n) {
  int64_t need = M->nrec + n;
  if (need > M->cap) {
    int64_t cap = std::max<int64_t>(need, 2 * M->cap);
    float* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, (size_t)cap * NREC * sizeof(float));
    if (e != hipSuccess) {
      mfail(LCEQ_ERR_HIP, "hipMalloc(%lld records): %s", (long long)cap, hipGetErrorString(e));
      return nullptr;
    }
    if (M->data) {
      e = hipMemcpy(d, M->data, (size_t)M->nrec * NREC * sizeof(float), hipMemcpyDeviceToDevice);
      if (e == hipSuccess) e = hipFree(M->data);
      if (e != hipSuccess) {
        (void)hipFree(d);
        mfail(LCEQ_ERR_HIP, "moving the training records: %s", hipGetErrorString(e));
        return nullptr;
      }
    }
    M->data = d;
    M->cap = cap;
  }
  if (hipStreamSynchronize(M->stream) != hipSuccess) {
    mfail(LCEQ_ERR_HIP, "model stream sync failed");
    return nullptr;
  }
  return M->data + (size_t)M->nrec * NREC;
}

void lceq::mlp::commit_records(lceq_mlp* M, int64_t n) { M->nrec += n; }

void lceq::mlp::launch_record(hipStream_t stream, const Geom& g, const double* U, const double* A,
                              int eo, int stride, int nrec, double ba, double b2, double b4,
                              float* out, int target) {
  int nb = (nrec + TB - 1) / TB;
  hipLaunchKernelGGL(k_mlp_record, dim3(nb), dim3(TB), 0, stream, g, U, A, eo, stride, nrec, ba,
                     b2, b4, out, target);
}

// Training pairs from the lattice's current configuration: for each parity,
// the features at the current F, then nsteps ahiggs1 Metropolis steps of
// that parity at the lattice's current step size (F does not change while
// one parity is updated), then the new a in the same frame; every stride-th
// site.  The Metropolis steps are kept (each is a valid update of the
// chain).  The Metropolis step rather than the ahiggs2 heatbath: the
// heatbath's Gaussian width beta_2 is negative in the broken phase.
extern "C" int lceq_mlp_collect(lceq_lattice* L, lceq_mlp* M, int nsteps, int stride) {
  if (!L || !M) return mfail(LCEQ_ERR_ARG, "null argument");
  if (nsteps < 1 || stride < 1) return mfail(LCEQ_ERR_ARG, "nsteps and stride must be >= 1");
  if (L->rng_mode != LCEQ_RNG_PHILOX) return mfail(LCEQ_ERR_ARG, "collect needs Philox streams");
  int nrec = (L->g.n + stride - 1) / stride;
  float* base = reserve_records(M, 2 * (int64_t)nrec);
  if (!base) return LCEQ_ERR_HIP;
  for (int eo = 0; eo < 2; ++eo) {
    float* out = base + (size_t)eo * nrec * NREC;
    launch_record(L->stream, L->g, L->U, L->A, eo, stride, nrec, L->beta_a, L->beta_2, L->beta_4,
                  out, 0);
    for (int h = 0; h < nsteps; ++h) launch_higgs(L, 1, eo, 0, nullptr, next_pass(L));
    launch_record(L->stream, L->g, L->U, L->A, eo, stride, nrec, L->beta_a, L->beta_2, L->beta_4,
                  out, 1);
  }
  int rc = check_launch();
  if (rc) return rc;
  MHIP(hipStreamSynchronize(L->stream));
  commit_records(M, 2 * (int64_t)nrec);
  return LCEQ_OK;
}

// Feature normalisation from the collected records (mean, 1/std; constant
// features get scale 1)
static int mlp_fit_norm(lceq_mlp* M) {
  std::vector<float> d((size_t)M->nrec * NREC);
  MHIP(hipMemcpy(d.data(), M->data, d.size() * sizeof(float), hipMemcpyDeviceToHost));
  for (int k = 0; k < NIN; ++k) {
    double s = 0.0, s2 = 0.0;
    for (int64_t i = 0; i < M->nrec; ++i) {
      double v = d[(size_t)i * NREC + k];
      s += v;
      s2 += v * v;
    }
    double mean = s / (double)M->nrec, var = s2 / (double)M->nrec - mean * mean;
    double sd = var > 1e-12 ? std::sqrt(var) : 0.0;
    M->host[NPARAM + k] = (float)mean;
    M->host[NPARAM + NIN + k] = sd > 1e-6 ? (float)(1.0 / sd) : 1.0f;
  }
  return LCEQ_OK;
}

// Adam on the collected records: `steps` mini-batches of `batch` records,
// each batch a stretch of a seeded affine permutation of the records.
// loss (optional, steps values) receives the mean negative log-likelihood
// per step (up to the constant 3/2 log 2 pi).
extern "C" int lceq_mlp_train(lceq_mlp* M, int steps, int batch, double lr, uint64_t seed,
                              int fit_norm, double* loss) {
  if (!M || steps < 0 || batch < TR_TB || batch % TR_TB)
    return mfail(LCEQ_ERR_ARG, "batch must be a positive multiple of %d", TR_TB);
  if (M->nrec < batch) return mfail(LCEQ_ERR_ARG, "only %lld records", (long long)M->nrec);
  int rc;
  if (fit_norm) {
    if ((rc = mlp_fit_norm(M))) return rc;
    if ((rc = mlp_upload(M))) return rc;
  }
  int nblk = batch / TR_TB;
  if (nblk > M->gpart_blocks) {
    if (M->gpart) MHIP(hipFree(M->gpart));
    if (M->lpart) MHIP(hipFree(M->lpart));
    M->gpart = nullptr;
    M->lpart = nullptr;
    MHIP(hipMalloc((void**)&M->gpart, (size_t)nblk * NPARAM * sizeof(float)));
    MHIP(hipMalloc((void**)&M->lpart, (size_t)nblk * sizeof(float)));
    M->gpart_blocks = nblk;
  }
  if (batch > M->idx_cap) {
    if (M->idx) MHIP(hipFree(M->idx));
    M->idx = nullptr;
    MHIP(hipMalloc((void**)&M->idx, (size_t)batch * sizeof(int)));
    M->idx_cap = batch;
  }
  if (M->adam_t == 0) {
    MHIP(hipMemsetAsync(M->adam_m, 0, NPARAM * sizeof(float), M->stream));
    MHIP(hipMemsetAsync(M->adam_v, 0, NPARAM * sizeof(float), M->stream));
  }
  const int64_t N = M->nrec;
  uint64_t st = seed;
  int64_t mult = 1, offs = 0, pos = N;
  std::vector<int> hidx(batch);
  std::vector<float> lp(nblk);
  for (int it = 0; it < steps; ++it) {
    if (pos + batch > N) {  // new epoch: a fresh affine permutation i -> (mult i + offs) mod N
      do {
        mult = (int64_t)(splitmix(st) % (uint64_t)N) | 1;
      } while (std::gcd(mult, N) != 1);
      offs = (int64_t)(splitmix(st) % (uint64_t)N);
      pos = 0;
    }
    for (int b = 0; b < batch; ++b)
      hidx[b] = (int)(((unsigned __int128)(mult) * (unsigned __int128)(pos + b) + offs) % N);
    pos += batch;
    MHIP(hipMemcpyAsync(M->idx, hidx.data(), batch * sizeof(int), hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(TR_TB), 0, M->stream, M->params, M->norm,
                       M->data, M->idx, batch, M->gpart, M->lpart);
    ++M->adam_t;
    float c1 = 1.0f - std::pow(0.9f, (float)M->adam_t), c2 = 1.0f - std::pow(0.999f, (float)M->adam_t);
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, M->stream, M->params,
                       M->adam_m, M->adam_v, M->gpart, nblk, (float)lr, c1, c2);
    if ((rc = check_launch())) return rc;
    if (loss) {
      MHIP(hipMemcpyAsync(lp.data(), M->lpart, nblk * sizeof(float), hipMemcpyDeviceToHost,
                          M->stream));
      MHIP(hipStreamSynchronize(M->stream));
      double l = 0.0;
      for (int b = 0; b < nblk; ++b) l += lp[b];
      loss[it] = l / batch;
    }
  }
  MHIP(hipMemcpyAsync(M->host.data(), M->params, NPARAM * sizeof(float), hipMemcpyDeviceToHost,
                      M->stream));
  MHIP(hipStreamSynchronize(M->stream));
  return mlp_upload(M);
}

extern "C" int lceq_mlp_set_sites_per_wave(lceq_mlp* M, int sites) {
  if (!M) return mfail(LCEQ_ERR_ARG, "null argument");
  if (sites != 0 && sites != 32 && sites != 64)
    return mfail(LCEQ_ERR_ARG, "sites per wave must be 0 (by lattice size), 32 or 64, not %d",
                 sites);
  M->sites_per_wave = sites;
  return LCEQ_OK;
}

extern "C" int lceq_mlp_eval(lceq_mlp* M, const float* x, int n, float* y) {
  if (!M || n < 1 || !x || !y) return mfail(LCEQ_ERR_ARG, "bad arguments");
  float *dx = nullptr, *dy = nullptr;
  MHIP(hipMalloc((void**)&dx, (size_t)n * NIN * sizeof(float)));
  if (hipMalloc((void**)&dy, (size_t)n * NOUT * sizeof(float)) != hipSuccess) {
    (void)hipFree(dx);
    return mfail(LCEQ_ERR_HIP, "hipMalloc failed");
  }
  int rc = LCEQ_OK;
  if (hipMemcpyAsync(dx, x, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice, M->stream) !=
      hipSuccess)
    rc = mfail(LCEQ_ERR_HIP, "copy failed");
  if (!rc) {
    hipLaunchKernelGGL(k_mlp_eval, dim3((n + TB - 1) / TB), dim3(TB), 0, M->stream, M->pack, dx, dy,
                       n);
    rc = check_launch();
  }
  if (!rc && (hipMemcpyAsync(y, dy, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost,
                             M->stream) != hipSuccess ||
              hipStreamSynchronize(M->stream) != hipSuccess))
    rc = mfail(LCEQ_ERR_HIP, "copy failed");
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

// both parities with the learned proposal; partials (always measured) in the
// Higgs slots
// Sites per wave of the learned Higgs launch: 64 (four MFMA column tiles,
// the default) or 32 (two tiles, half the lanes idle in the per-site FP64
// work).  32 shortens one wave's network evaluations but not its FP64
// chain; at 48^3 (fewer waves than SIMDs) it measured 15.6 us per parity
// against 13.8 us for 64.  The fields do not depend on the choice.
static int mlp_sites_per_wave(const lceq_lattice*, const lceq_mlp* M) {
  return M->sites_per_wave ? M->sites_per_wave : 64;
}
static int mlp_blocks(const lceq_lattice* L, const lceq_mlp* M) {
  int per_block = (TB / 64) * mlp_sites_per_wave(L, M);
  return (L->g.n + per_block - 1) / per_block;
}

static int mlp_higgs_sweep(lceq_lattice* L, lceq_mlp* M) {
  if (L->rng_mode != LCEQ_RNG_PHILOX) return mfail(LCEQ_ERR_ARG, "the learned proposal draws Philox");
  MlpArgs a;
  a.g = L->g;
  a.U = L->U;
  a.A = L->A;
  a.beta_a = L->beta_a;
  a.beta_2 = L->beta_2;
  a.beta_4 = L->beta_4;
  a.key = Philox{(uint32_t)L->seed, (uint32_t)(L->seed >> 32)};
  a.pack = M->pack;
  a.pstride = L->pstride;
  int spw = mlp_sites_per_wave(L, M);
  a.nb = mlp_blocks(L, M);
  prof_begin(L, KC_HIGGS);
  for (int eo = 0; eo < 2; ++eo) {
    a.eo = eo;
    a.pass = next_pass(L);
    a.part = L->partials + (size_t)(6 + eo) * LCEQ_NQ * L->pstride;
    prof_count(L, KC_HIGGS);
    if (spw == 32)
      hipLaunchKernelGGL(k_higgs_mlp<2>, dim3(a.nb), dim3(TB), 0, L->stream, a);
    else
      hipLaunchKernelGGL(k_higgs_mlp<4>, dim3(a.nb), dim3(TB), 0, L->stream, a);
  }
  prof_end(L, KC_HIGGS);
  return check_launch();
}

extern "C" int lceq_higgs_mlp(lceq_lattice* L, lceq_mlp* M, double* accept) {
  if (!L || !M) return mfail(LCEQ_ERR_ARG, "null argument");
  int rc = mlp_higgs_sweep(L, M);
  if (rc) return rc;
  launch_finalize(L, 0, 0, 2, 1, nullptr, mlp_blocks(L, M));
  if (accept) {
    double o[LCEQ_OBS_N];
    if ((rc = read_obs(L, o))) return rc;
    *accept = o[OBS_ACC_HIGGS];
  }
  return LCEQ_OK;
}

// nsweep sweeps of 1 KP + n_or overrelaxation link sweeps + the learned
// Higgs update; with measflag one history row per sweep (as lceq_sweep)
extern "C" int lceq_sweep_mlp(lceq_lattice* L, lceq_mlp* M, int nsweep, int n_or, int measflag) {
  if (!L || !M) return mfail(LCEQ_ERR_ARG, "null argument");
  if (nsweep < 0 || n_or < 0) return mfail(LCEQ_ERR_ARG, "nsweep, n_or must be >= 0");
  if (measflag && L->hist_rows + nsweep > L->hist_cap)
    return mfail(LCEQ_ERR_ARG, "history full (%d rows); call lceq_history", L->hist_cap);
  int rc;
  for (int it = 0; it < nsweep; ++it) {
    double* row = measflag ? L->hist + (size_t)(L->hist_rows++) * LCEQ_OBS_N : nullptr;
    int meas_kp = measflag && n_or == 0;
    if ((rc = link_sweep(L, 1, meas_kp, 1, meas_kp))) return rc;
    for (int k = 0; k < n_or; ++k) {
      int m = measflag && k == n_or - 1;
      if ((rc = link_sweep(L, 0, m, 1, m))) return rc;
    }
    if ((rc = mlp_higgs_sweep(L, M))) return rc;
    launch_finalize(L, measflag ? (n_or ? 2 : 1) : 0, measflag, 2, 1, row, mlp_blocks(L, M));
  }
  return check_launch();
}
