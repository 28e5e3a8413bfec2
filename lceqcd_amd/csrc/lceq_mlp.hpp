// lceq_mlp.hpp -- device code of the learned proposal for the adjoint-Higgs
// update, shared by the single-lattice launch (lceq_mlp.hip) and the
// replica batch (lceq_batch.hip): network layout, frame, bf16 MFMA
// inference, the Gaussian proposal and one parity's per-site update.  No
// counterpart in the reference; the plain update it competes with is the
// Metropolis step of f/ahiggs1.f.
#pragma once
#include <math.h>
#include <stdint.h>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"
#include "lceq_sites.hpp"

namespace lceq {
namespace mlp {

// sites per block of the learned Higgs launches: four waves of 64 sites
constexpr int MLP_TB = 256;

constexpr int NIN = 10, NH = 64, NOUT = 6;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NIN * NH, OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NOUT;
constexpr int NPARAM = OFF_B3 + NOUT;  // 5254
static_assert(NPARAM == LCEQ_MLP_NPARAM, "parameter count");
constexpr int NSTATE = NPARAM + 2 * NIN;  // + feature mean and inverse scale
static_assert(NSTATE == LCEQ_MLP_NSTATE, "state size");
constexpr int NREC = NIN + 3;             // training record: features, target y
constexpr float S_MIN = -7.0f, S_MAX = 4.0f;  // clamp of the log-scales

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
// update always measures), one row entry per logical block blk of 4 waves x
// 16 ST sites.  Every lane of the block must call it (MFMA and block sums).
template <int ST>
__device__ __forceinline__ void mlp_higgs_site(const MlpArgs& p, int blk, MlpLds& S) {
  mlp_stage_bias(p.pack, S);
  // 16 ST sites per wave, lane l < 16 ST owns site (wave) 16 ST + l
  int lane = threadIdx.x & 63;
  unsigned s = (unsigned)((blk * (MLP_TB / 64) + (threadIdx.x >> 6)) * 16 * ST + lane);
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


// ---- host helpers of lceq_mlp.hip used by the replica batch

// the inference weights on the device (valid until the model is destroyed;
// retraining rewrites them in place)
const MlpPack* device_pack(const lceq_mlp* M);
// device room for n more training records (NREC floats each) after the ones
// collected so far; nullptr with the error set if it cannot be allocated.
// The records count once commit_records(M, n) is called.
float* reserve_records(lceq_mlp* M, int64_t n);
void commit_records(lceq_mlp* M, int64_t n);
// one parity's records (every stride-th site, nrec of them) on `stream`:
// target = 0 writes the inputs at the current F, target = 1 the current a in
// the same frame
void launch_record(hipStream_t stream, const Geom& g, const double* U, const double* A, int eo,
                   int stride, int nrec, double ba, double b2, double b4, float* out, int target);

}  // namespace mlp
}  // namespace lceq
