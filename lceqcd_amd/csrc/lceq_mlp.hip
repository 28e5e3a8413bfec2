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

namespace {

constexpr int NIN = 10, NH = 64, NOUT = 6;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NOUT;
constexpr int NPARAM = OFF_B3 + NOUT;  // 5254
static_assert(NPARAM == LCEQ_MLP_NPARAM, "parameter count");
constexpr int NSTATE = NPARAM + 2 * NIN;  // + feature mean and inverse scale
static_assert(NSTATE == LCEQ_MLP_NSTATE, "state size");
constexpr int NREC = NIN + 3;             // training record: features, target y
constexpr float S_MIN = -7.0f, S_MAX = 4.0f;  // clamp of the log-scales
constexpr int TB = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx8 __attribute__((ext_vector_type(8)));

// Inference weights as A operands of v_mfma_f32_16x16x32_bf16 (M = 16
// output units x K = 32 inputs): lane l holds row l % 16 (an output unit) and
// K elements 8 (l / 16) .. + 7.  The sites are the N (column) dimension, so a
// layer's result (D: lane l holds column l % 16, rows 4 (l / 16) .. + 3) is,
// after bias and SiLU, the next layer's B operand without leaving the lane:
// K element 8g + j of block kb is hidden unit 32 kb + 4g + j (j < 4) or
// 32 kb + 16 + 4g + (j - 4) (j >= 4), the rows that D tiles 2 kb and
// 2 kb + 1 give lane group g.  Layers 2 and 3 are packed in that K order.
struct MlpPack {
  bf16x8 w1[4][64];     // [n tile][lane], K = 10 inputs padded to 32
  bf16x8 w2[4][2][64];  // [n tile][k block][lane], K in the order above
  bf16x8 w3[2][64];     // [k block][lane], M = 6 outputs padded to 16
  float b1[NH], b2[NH], b3[16];
  float xmean[NIN], xinv[NIN];
};
// hidden unit of K element 8 g + j of block kb (layers 2 and 3)
__host__ __device__ constexpr int hidden_of_k(int kb, int g, int j) {
  return 32 * kb + (j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4));
}

// ------------------------------------------------------------- frame + features

struct Frame {
  double e[3][3];  // e[i] = axis i
  double fn;       // |F|
};

__host__ __device__ inline Frame make_frame(const double F[3]) {
  Frame f;
  f.fn = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  double e1[3] = {1.0, 0.0, 0.0};
  if (f.fn > 0.0)
    for (int i = 0; i < 3; ++i) e1[i] = F[i] / f.fn;
  double h[3] = {0.0, 0.0, 1.0};
  if (fabs(e1[2]) >= 0.9) {
    h[0] = 1.0;
    h[2] = 0.0;
  }
  double e2[3] = {h[1] * e1[2] - h[2] * e1[1], h[2] * e1[0] - h[0] * e1[2],
                  h[0] * e1[1] - h[1] * e1[0]};
  double n2 = sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
  for (int i = 0; i < 3; ++i) e2[i] = e2[i] / n2;
  double e3[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                  e1[0] * e2[1] - e1[1] * e2[0]};
  for (int i = 0; i < 3; ++i) {
    f.e[0][i] = e1[i];
    f.e[1][i] = e2[i];
    f.e[2][i] = e3[i];
  }
  return f;
}
__host__ __device__ inline void to_frame(const Frame& f, const double a[3], double y[3]) {
  for (int i = 0; i < 3; ++i) y[i] = (f.e[i][0] * a[0] + f.e[i][1] * a[1]) + f.e[i][2] * a[2];
}
__host__ __device__ inline void from_frame(const Frame& f, const double y[3], double a[3]) {
  for (int i = 0; i < 3; ++i) a[i] = (f.e[0][i] * y[0] + f.e[1][i] * y[1]) + f.e[2][i] * y[2];
}
// raw inputs: a in the frame, the hopping sum in the frame (|F|, 0, 0), |F|,
// beta_A, beta_2, beta_4
__host__ __device__ inline void raw_features(const double y[3], double fn, double ba, double b2,
                                             double b4, float x[NIN]) {
  x[0] = (float)y[0];
  x[1] = (float)y[1];
  x[2] = (float)y[2];
  x[3] = (float)fn;
  x[4] = 0.0f;
  x[5] = 0.0f;
  x[6] = (float)fn;
  x[7] = (float)ba;
  x[8] = (float)b2;
  x[9] = (float)b4;
}

// ----------------------------------------------------------------- inference

// Per-wave scratch of one evaluation: the sites' bf16 input rows (K padded
// to 16), the fp32 outputs, and the biases in D-row order, the accumulator
// operand of each MFMA chain: bias[4 nt + g] = units 16 nt + 4g .. + 3 of
// b1, then b2 from 16, then b3 from 32
struct alignas(16) MlpLds {
  bf16x8 x[64][2];
  float o[64][8];
  floatx4 bias[36];
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
// inference SiLU: hardware exp and reciprocal (the activations are rounded
// to bf16 next; exactness of the chain only needs a deterministic network)
__device__ __forceinline__ float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
// Order this wave's LDS accesses: every lane of a wave writes and reads only
// its own wave's scratch, and one wave's LDS operations complete in program
// order, so the evaluation needs no block barrier.
__device__ __forceinline__ void wave_sync() { __builtin_amdgcn_wave_barrier(); }

// This lane's share of the inference weights (A-operand fragments), held
// in registers for the whole kernel: read from LDS per tile instead, they
// cost more in latency than the registers they free
struct MlpRegs {
  bf16x8 w1[4], w2[4][2], w3[2];
};

__device__ __forceinline__ void mlp_load(const MlpPack* __restrict__ P, MlpRegs& R) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    R.w1[nt] = P->w1[nt][l];
    R.w2[nt][0] = P->w2[nt][0][l];
    R.w2[nt][1] = P->w2[nt][1][l];
  }
  R.w3[0] = P->w3[0][l];
  R.w3[1] = P->w3[1][l];
}

// The biases into this wave's scratch (once per kernel, before mlp_eval)
__device__ __forceinline__ void mlp_stage_bias(const MlpPack* __restrict__ P, MlpLds& S) {
  const int l = threadIdx.x & 63;
  if (l < 16)
    S.bias[l] = *reinterpret_cast<const floatx4*>(&P->b1[4 * l]);
  else if (l < 32)
    S.bias[l] = *reinterpret_cast<const floatx4*>(&P->b2[4 * (l - 16)]);
  else if (l < 36)
    S.bias[l] = *reinterpret_cast<const floatx4*>(&P->b3[4 * (l - 32)]);
  wave_sync();
}

// SiLU of two D tiles (units 32 kb + 4g + i and 32 kb + 16 + 4g + i, bias
// included) -> the bf16 B operand of K block kb
__device__ __forceinline__ bf16x8 silu_pack(const floatx4& d0, const floatx4& d1) {
  floatx8 v;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[i] = silu_fast(d0[i]);
    v[4 + i] = silu_fast(d1[i]);
  }
  return __builtin_convertvector(v, bf16x8);
}

// One evaluation for the 16 ST sites of this wave (lane l < 16 ST = site;
// the other lanes' outputs are zero).  Every lane of the wave must call it.
// The sites are the MFMA columns: ST tiles of 16.
template <int ST>
__device__ void mlp_eval(const MlpPack* __restrict__ P, const MlpRegs& R, const float (&xin)[NIN],
                         float (&out)[NOUT], MlpLds& S) {
  const int l = threadIdx.x & 63, c = l & 15, g = l >> 4;
  // this site's inputs, normalised, as a bf16 row of 16 (10 used)
  floatx8 f0, f1;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    f0[k] = (xin[k] - P->xmean[k]) * P->xinv[k];
    f1[k] = 8 + k < NIN ? (xin[8 + k] - P->xmean[8 + k]) * P->xinv[8 + k] : 0.0f;
  }
  if (l < 16 * ST) {
    S.x[l][0] = __builtin_convertvector(f0, bf16x8);
    S.x[l][1] = __builtin_convertvector(f1, bf16x8);
  }
  wave_sync();
  const bf16x8 zero8 = {};
  // one site tile (16 columns) at a time through the three layers: the
  // tiles are independent, and this keeps one tile's activations live
#pragma unroll
  for (int st = 0; st < ST; ++st) {
    // B operand of layer 1: site 16 st + c, inputs 8g .. 8g + 7 (zero for g >= 2)
    bf16x8 xb = g < 2 ? S.x[16 * st + c][g] : zero8;
    // layer 1: D[unit][site] = W1^T [64 x 32] X^T [32 x 16]
    floatx4 acc[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[nt], xb, S.bias[4 * nt + g], 0, 0, 0);
    bf16x8 h0 = silu_pack(acc[0], acc[1]);
    bf16x8 h1 = silu_pack(acc[2], acc[3]);
    // layer 2: [64 x 64] x [64 x 16]
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      floatx4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[nt][0], h0, S.bias[16 + 4 * nt + g],
                                                          0, 0, 0);
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[nt][1], h1, t, 0, 0, 0);
    }
    h0 = silu_pack(acc[0], acc[1]);
    h1 = silu_pack(acc[2], acc[3]);
    // layer 3: [16 (6 used) x 64] x [64 x 16]; lane groups 0 and 1 hold the
    // outputs 0-3 and 4-7 of site 16 st + c
    floatx4 t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w3[0], h0, S.bias[32 + g], 0, 0, 0);
    t = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w3[1], h1, t, 0, 0, 0);
    if (g < 2) *reinterpret_cast<floatx4*>(&S.o[16 * st + c][4 * g]) = t;
  }
  wave_sync();
#pragma unroll
  for (int j = 0; j < NOUT; ++j) out[j] = l < 16 * ST ? S.o[l][j] : 0.0f;
  wave_sync();
}

// Gaussian proposal of the network outputs: mean and scale in the frame
__device__ __forceinline__ void proposal(const float (&o)[NOUT], double mu[3], double sig[3],
                                         double ls[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    mu[i] = (double)o[i];
    ls[i] = (double)fminf(fmaxf(o[3 + i], S_MIN), S_MAX);
    sig[i] = exp(ls[i]);
  }
}
// log q up to the constant -3/2 log(2 pi), which cancels in the ratio
__device__ __forceinline__ double log_q(const double y[3], const double mu[3], const double sig[3],
                                        const double ls[3]) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double t = (y[i] - mu[i]) / sig[i];
    r = r - ls[i] - 0.5 * (t * t);
  }
  return r;
}

struct MlpArgs {
  Geom g;
  const double* U;
  double* A;
  int eo;
  double beta_a, beta_2, beta_4;
  Philox key;
  uint64_t pass;
  const MlpPack* pack;
  double* part;
  int pstride, nb;
};

// One Higgs parity with the learned proposal.  Uniforms of site s (Philox
// stream 0 of this pass): k = 0, 1 and 2, 3 give the three Gaussians by
// Box-Muller, k = 4 the accept.  Partials: accepts, rho, rho^2, a.F (the
// update always measures).
template <int ST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_higgs_mlp(MlpArgs p) {
  __shared__ MlpLds lds[4];
  MlpLds& S = lds[threadIdx.x >> 6];
  mlp_stage_bias(p.pack, S);
  int blk = xcd_block(p.nb);
  // 16 ST sites per wave, lane l < 16 ST owns site (wave) 16 ST + l
  int lane = threadIdx.x & 63;
  unsigned s = (unsigned)((blk * (TB / 64) + (threadIdx.x >> 6)) * 16 * ST + lane);
  bool valid = lane < 16 * ST && s < (unsigned)p.g.n;
  unsigned sl = valid ? s : 0u;
  const Geom& g = p.g;
  Coord c = site_coord(g, p.eo, (int)sl);
  ForceLinks k;
  load_force_links(p.U, p.A, g, p.eo, sl, c, k);
  vec3 av = ld_higgs(p.A, g, p.eo, sl);
  __builtin_amdgcn_sched_barrier(0);
  // under the load latency: the uniforms and the Box-Muller Gaussians
  double z[3], racc;
  {
    double u0, u1, u2, u3, dummy;
    philox_pair(p.key, sl, 0, p.pass, 0, u0, u1);
    philox_pair(p.key, sl, 0, p.pass, 1, u2, u3);
    philox_pair(p.key, sl, 0, p.pass, 2, racc, dummy);
    double r01 = sqrt(-2.0 * log(u0)), r23 = sqrt(-2.0 * log(u2));
    double s1, c1;
    sincospi(2.0 * u1, &s1, &c1);  // sin, cos(2 pi u1) without a rounded 2 pi
    z[0] = r01 * c1;
    z[1] = r01 * s1;
    z[2] = r23 * cospi(2.0 * u3);
  }
  vec3 F = force_sum(k);
  Frame fr = make_frame(F.x);
  // site action terms of the current a, f/ahiggs1.f:64-92
  double sfo = 0.0, ro = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sfo = sfo + av.x[i] * F.x[i];
    ro = ro + av.x[i] * av.x[i];
  }
  double ya[3], fn = fr.fn;
  to_frame(fr, av.x, ya);
  float xin[NIN], o[NOUT];
  raw_features(ya, fn, p.beta_a, p.beta_2, p.beta_4, xin);
  MlpRegs W;
  mlp_load(p.pack, W);
  mlp_eval<ST>(p.pack, W, xin, o, S);
  // forward proposal a' and log q(a' | a); only these and the action terms
  // stay live across the second evaluation
  double an[3], ypr[3], lqf, sfn = 0.0, rn = 0.0;
  {
    double mu[3], sig[3], ls[3], yp[3];
    proposal(o, mu, sig, ls);
#pragma unroll
    for (int i = 0; i < 3; ++i) yp[i] = mu[i] + sig[i] * z[i];
    from_frame(fr, yp, an);
    to_frame(fr, an, ypr);
    lqf = log_q(ypr, mu, sig, ls);
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    sfn = sfn + an[i] * F.x[i];
    rn = rn + an[i] * an[i];
  }
  raw_features(ypr, fn, p.beta_a, p.beta_2, p.beta_4, xin);
  mlp_eval<ST>(p.pack, W, xin, o, S);
  double lqr;
  {
    double mu2[3], sig2[3], ls2[3];
    proposal(o, mu2, sig2, ls2);
    lqr = log_q(ya, mu2, sig2, ls2);
  }
  double boltz = exp(-(p.beta_a * (sfo - sfn)) + p.beta_2 * (ro - rn) +
                     p.beta_4 * (ro * ro - rn * rn) + (lqr - lqf));
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  double af = sfo;  // a.F of the site after the update
  if (boltz >= racc) {
    if (valid) st_higgs(p.A, g, p.eo, s, vec3{{an[0], an[1], an[2]}});
    ro = rn;
    af = sfn;
    red[0] = 1.0;
  }
  red[1] = ro;
  red[2] = ro * ro;
  red[3] = af;
  if (!valid) red[0] = red[1] = red[2] = red[3] = 0.0;
  block_sum_store<4>(red, p.part, p.pstride, blk);
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
extern "C" int lceq_mlp_collect(lceq_lattice* L, lceq_mlp* M, int nsteps, int stride) {
  if (!L || !M) return mfail(LCEQ_ERR_ARG, "null argument");
  if (nsteps < 1 || stride < 1) return mfail(LCEQ_ERR_ARG, "nsteps and stride must be >= 1");
  if (L->rng_mode != LCEQ_RNG_PHILOX) return mfail(LCEQ_ERR_ARG, "collect needs Philox streams");
  int nrec = (L->g.n + stride - 1) / stride;
  int64_t need = M->nrec + 2 * (int64_t)nrec;
  if (need > M->cap) {
    int64_t cap = std::max<int64_t>(need, 2 * M->cap);
    float* d = nullptr;
    MHIP(hipMalloc((void**)&d, (size_t)cap * NREC * sizeof(float)));
    if (M->data) {
      MHIP(hipMemcpy(d, M->data, (size_t)M->nrec * NREC * sizeof(float), hipMemcpyDeviceToDevice));
      MHIP(hipFree(M->data));
    }
    M->data = d;
    M->cap = cap;
  }
  MHIP(hipStreamSynchronize(M->stream));
  for (int eo = 0; eo < 2; ++eo) {
    float* out = M->data + (size_t)(M->nrec + (int64_t)eo * nrec) * NREC;
    int nb = (nrec + TB - 1) / TB;
    hipLaunchKernelGGL(k_mlp_record, dim3(nb), dim3(TB), 0, L->stream, L->g, L->U, L->A, eo, stride,
                       nrec, L->beta_a, L->beta_2, L->beta_4, out, 0);
    for (int h = 0; h < nsteps; ++h) launch_higgs(L, 1, eo, 0, nullptr, next_pass(L));
    hipLaunchKernelGGL(k_mlp_record, dim3(nb), dim3(TB), 0, L->stream, L->g, L->U, L->A, eo, stride,
                       nrec, L->beta_a, L->beta_2, L->beta_4, out, 1);
  }
  int rc = check_launch();
  if (rc) return rc;
  MHIP(hipStreamSynchronize(L->stream));
  M->nrec += 2 * (int64_t)nrec;
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
