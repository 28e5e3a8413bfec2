// lceq_batch.hip -- many independent lattices of uneven sizes and couplings
// (a phase-diagram scan) swept together: one launch per (parity, mu) link
// class and per Higgs parity covers every replica of a replica group, so the
// device sees the sum of the group's sites per launch.  A batch larger than
// the 256 MB Infinity Cache is split into groups of about 288 MiB of fields,
// and a sweep runs its launches group after group: the 33 launches of a
// group's sweep then find its fields in the Infinity Cache instead of HBM
// (the 539 MB, 96-replica scan sweeps 25% faster as two groups than as one),
// at the cost of one more round of launches per group.  Eight sweeps'
// launches are captured once in a hipGraph and replayed; the Philox pass
// number comes from a device-side pass counter (with the sweep counter that
// indexes the history rows) so that the captured graph stays valid from sweep
// to sweep.  The launch shape (threads per block, XCD-aware block order) is a
// runtime choice that lceq_batch_autotune picks by timing.
//
// Replica r with seed s_r follows exactly the trajectory that a single
// lceq_lattice with seed s_r, started at pass 0, follows under
// lceq_sweep(LCEQ_SWEEP_KP_OR_METRO): the per-site code (lceq_sites.hpp),
// the Philox counters and the reduction order are the same.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"
#include "lceq_mlp.hpp"
#include "lceq_sites.hpp"

using namespace lceq;

namespace {

constexpr int TB_MIN = 64;  // smallest block of the launch shapes
constexpr int FIN_TB = 1024;

struct Replica {
  Geom g;
  double *U, *A;
  double beta, beta_a, beta_2, beta_4;
  Philox key;
  int blk0, nb;     // first global block, number of blocks
  int mblk0, mnb;   // the same for the learned Higgs launches (256 sites per block)
  double* part;     // [LCEQ_NSLOT][LCEQ_NQ][pstride]
  int pstride;
  double* obs;      // [LCEQ_OBS_N]
};

struct BatchCommon {
  const Replica* reps;
  const int* blk2rep;
  int total_nb;           // blocks of this launch (one replica group's)
  int blk_base;           // global index of the launch's first block
  // device counters: ctr[0] sweeps (history row), ctr[1] the Philox pass
  // of the first launch of the current sweep; launch j of a sweep uses pass
  // ctr[1] + j, as the j-th next_pass() of a single lattice's sweep
  const uint64_t* ctr;
  int xcd;                // XCD-aware block order (else launch order)
};

__device__ __forceinline__ int batch_block(const BatchCommon& c, int& r, const Replica*& R) {
  int gblk = c.blk_base + (c.xcd ? xcd_block(c.total_nb) : (int)blockIdx.x);
  r = c.blk2rep[gblk];
  R = c.reps + r;
  return gblk - R->blk0;
}

// One (eo, mu) link class of every replica, TBX sites per block
template <bool KP, int TBX>
__global__ __launch_bounds__(TBX) void k_link_batch(BatchCommon c, int eo, int mu, int pass_in_sweep,
                                                    int measflag, int reduce) {
  int r;
  const Replica* R;
  int lblk = batch_block(c, r, R);
  unsigned s = (unsigned)(lblk * TBX + threadIdx.x);
  double red[3] = {0.0, 0.0, 0.0};
  if (s < (unsigned)R->g.n) {
    LinkSite p;
    p.g = R->g;
    p.U = R->U;
    p.A = R->A;
    p.eo = eo;
    p.mu = mu;
    p.beta = R->beta;
    p.beta_a = R->beta_a;
    p.fed = nullptr;
    p.key = R->key;
    p.pass = c.ctr[1] + (uint64_t)pass_in_sweep;
    p.measflag = measflag;
    p.renorm = 1;
    link_update_site<KP, false>(p, s, red);
  }
  if (reduce)
    block_sum_store<3, TBX / 64>(red, R->part + (size_t)(eo * 3 + mu) * LCEQ_NQ * R->pstride,
                                 R->pstride, lblk);
}

template <int TBX>
__global__ __launch_bounds__(TBX) void k_higgs_batch(BatchCommon c, int eo, int pass_in_sweep,
                                                     int measflag) {
  int r;
  const Replica* R;
  int lblk = batch_block(c, r, R);
  unsigned s = (unsigned)(lblk * TBX + threadIdx.x);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  if (s < (unsigned)R->g.n) {
    HiggsSite p;
    p.g = R->g;
    p.U = R->U;
    p.A = R->A;
    p.eo = eo;
    p.imode = 0;
    p.beta_a = R->beta_a;
    p.beta_2 = R->beta_2;
    p.beta_4 = R->beta_4;
    p.delta = R->obs + OBS_DELTA;
    p.fed = nullptr;
    p.key = R->key;
    p.pass = c.ctr[1] + (uint64_t)pass_in_sweep;
    p.measflag = measflag;
    higgs_update_site<1, false>(p, s, red);
  }
  block_sum_store<4, TBX / 64>(red, R->part + (size_t)(6 + eo) * LCEQ_NQ * R->pstride, R->pstride,
                               lblk);
}

// One Higgs parity of every replica with the learned proposal: the per-site
// code of the single-lattice launch (mlp::mlp_higgs_site, 64 sites per wave,
// 256 per block) on the replica's fields, couplings, key and partial-sum
// slots, one network for all replicas (the couplings are network inputs).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_higgs_mlp_batch(
    BatchCommon c, const int* __restrict__ mblk2rep, int total_mnb, int mblk_base,
    const mlp::MlpPack* __restrict__ pack, int eo, int pass_in_sweep) {
  __shared__ mlp::MlpLds lds[4];
  int gblk = mblk_base + (c.xcd ? xcd_block(total_mnb) : (int)blockIdx.x);
  const Replica* R = c.reps + mblk2rep[gblk];
  mlp::MlpArgs a;
  a.g = R->g;
  a.U = R->U;
  a.A = R->A;
  a.eo = eo;
  a.beta_a = R->beta_a;
  a.beta_2 = R->beta_2;
  a.beta_4 = R->beta_4;
  a.key = R->key;
  a.pass = c.ctr[1] + (uint64_t)pass_in_sweep;
  a.pack = pack;
  a.part = R->part + (size_t)(6 + eo) * LCEQ_NQ * R->pstride;
  a.pstride = R->pstride;
  a.nb = R->mnb;
  mlp::mlp_higgs_site<4>(a, gblk - R->mblk0, lds[threadIdx.x >> 6]);
}

// Per-replica close of a sweep (block r): the same reduction and slot order
// as k_finalize of the single-lattice path, then the history row.
// higgs_kind 1: ahiggs1 (Higgs observables with meas, step adaptation);
// 2: the learned proposal (Higgs observables always, as lceq_sweep_mlp
// records them, no step adaptation; its partials come from R->mnb blocks).
__global__ __launch_bounds__(1024) void k_finalize_batch(const Replica* __restrict__ reps,
                                                         int rep_base, int link_kind, int meas,
                                                         int higgs_kind,
                                                         const uint64_t* __restrict__ ctr,
                                                         double* __restrict__ hist, int hist_cap,
                                                         int nrep) {
  constexpr int NV = 6 * 3 + 2 * 4;
  __shared__ double tot[32];
  const int r = rep_base + (int)blockIdx.x;
  const Replica* R = reps + r;
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int q = w + 16 * h;
    if (q >= NV) continue;
    bool on = q < 18 ? link_kind != 0 : true;
    double t = 0.0;
    if (on) {
      size_t row = q < 18 ? (size_t)(q / 3) * LCEQ_NQ + q % 3
                          : (size_t)(6 + (q - 18) / 4) * LCEQ_NQ + (q - 18) % 4;
      const double* x = R->part + row * R->pstride;
      int nbq = q < 18 || higgs_kind == 1 ? R->nb : R->mnb;
      for (int i0 = lane; i0 < nbq; i0 += 64 * 8) {
        double y[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) y[u] = i0 + 64 * u < nbq ? x[i0 + 64 * u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) t += y[u];
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    }
    if (lane == 0) tot[q] = t;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double* v = tot;
  double* obs = R->obs;
  double vol = 2.0 * (double)R->g.n;
  if (link_kind) {
    double plaq = 0.0, nacch = 0.0, acc1 = 0.0;
    for (int slot = 0; slot < 6; ++slot) {
      plaq = plaq + v[slot * 3];
      nacch += v[slot * 3 + 1];
      acc1 += v[slot * 3 + 2];
    }
    obs[OBS_NACCH] = nacch;
    obs[OBS_NACC1] = acc1;
    if (meas == 1) obs[OBS_PLAQ0] = plaq / (12.0 * vol);
    double accept;
    if (link_kind == 1) {
      accept = nacch;
      if (meas == 1) accept = (accept / (3.0 * vol)) * (acc1 / (12.0 * vol));
    } else {
      accept = acc1 / (3.0 * vol);
    }
    obs[OBS_ACC_LINK] = accept;
  }
  double acc = v[18] + v[22];
  double r2 = v[19] + v[23], r4 = v[20] + v[24], kin = v[21] + v[25];
  int hmeas = higgs_kind == 2 ? 1 : meas;
  if (hmeas == 1) {
    obs[OBS_A02] = r2 / vol;
    obs[OBS_A04] = r4 / vol;
    obs[OBS_A0KIN] = kin / (6.0 * vol);
  }
  double accept = acc / vol;
  obs[OBS_ACC_HIGGS] = accept;
  if (higgs_kind == 1) {
    double d = obs[OBS_DELTA];
    if (fabs(accept - 0.5) > 0.1) d = (accept > 0.5) ? 1.1131241 * d : 0.8989669 * d;
    obs[OBS_DELTA] = d;
  }
  if (meas == 1 && hist) {
    double* row = hist + ((size_t)(ctr[0] % (uint64_t)hist_cap) * nrep + r) * LCEQ_OBS_N;
    for (int i = 0; i < LCEQ_OBS_N; ++i) row[i] = obs[i];
  }
}

// advance the counters past a sweep (dsweep = 1) or past collect steps
// (dsweep = 0) that used dpass passes
__global__ void k_tick(uint64_t* ctr, int dsweep, int dpass) {
  if (threadIdx.x == 0) {
    ctr[0] += (uint64_t)dsweep;
    ctr[1] += (uint64_t)dpass;
  }
}

}  // namespace

#define bfail lceq::set_error

#define BHIP(call)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return bfail(LCEQ_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));       \
  } while (0)

struct lceq_batch;
static int batch_upload_table(lceq_batch* B);

// A run of consecutive replicas swept together: every launch of a sweep
// covers one group, and the groups of a sweep run one after another.
struct ReplicaGroup {
  int rep0, nrep;  // replicas [rep0, rep0 + nrep)
  int blk0, nb;    // their blocks of the update launches at the launch shape
  int mblk0, mnb;  // and of the learned Higgs launches
  int64_t bytes;   // their link and Higgs fields
};

struct lceq_batch {
  int nrep;
  std::vector<Replica> reps;  // host copy of the device table
  std::vector<int64_t> field_off;
  Replica* d_reps;
  int* d_blk2rep;
  int total_nb;
  int64_t total_sites;
  double *pool, *parts, *obs, *hist;
  size_t pool_doubles, parts_doubles;
  int hist_cap;
  uint64_t* d_ctr;       // [2]: sweeps, Philox pass base (BatchCommon::ctr)
  uint64_t sweeps_done;  // host mirror of the device counter
  hipStream_t stream;
  hipGraphExec_t graph;  // GRAPH_SWEEPS sweeps
  int graph_n_or, graph_meas;
  const mlp::MlpPack* graph_pack;  // the learned proposal's weights, or null (ahiggs1)
  int tb, xcd;           // launch shape
  // learned Higgs launches: 256-site blocks, independent of the launch shape
  int* d_mblk2rep;
  int total_mnb;
  // replica groups: fields of at most about group_bytes per group (0: one
  // group), so that a group stays in the Infinity Cache through its sweep
  int64_t group_bytes;
  std::vector<ReplicaGroup> groups;
};

// Default group size: the fields of a group about fill the 256 MB Infinity
// Cache (the 539 MB replica scan then runs as two groups of 270 MB).
constexpr int64_t DEFAULT_GROUP_BYTES = (int64_t)288 << 20;

constexpr int GRAPH_SWEEPS = 8;

static int batch_upload_table(lceq_batch* B) {
  BHIP(hipMemcpyAsync(B->d_reps, B->reps.data(), B->reps.size() * sizeof(Replica),
                      hipMemcpyHostToDevice, B->stream));
  BHIP(hipStreamSynchronize(B->stream));
  return LCEQ_OK;
}

// Split the replicas into ceil(total / group_bytes) runs of about equal
// field bytes (one run when group_bytes is 0 or the batch fits).
static void batch_groups(lceq_batch* B) {
  std::vector<int64_t> by(B->nrep);
  int64_t total = 0;
  for (int r = 0; r < B->nrep; ++r) total += by[r] = 30 * 8 * (int64_t)B->reps[r].g.npad;
  const int64_t gb = B->group_bytes;
  int64_t ng = gb > 0 ? total / gb + (total % gb != 0) : 1;
  ng = std::max<int64_t>(1, std::min<int64_t>(ng, B->nrep));
  B->groups.clear();
  int64_t cum = 0;
  int r0 = 0;
  for (int r = 0; r < B->nrep; ++r) {
    cum += by[r];
    int64_t g = (int64_t)B->groups.size();
    int64_t left = B->nrep - 1 - r, todo = ng - 1 - g;
    // close the group once it reaches its share of the total, or when the
    // replicas left are just enough for one per group still to come
    if (r == B->nrep - 1 ||
        (g < ng - 1 && left >= todo && (cum * ng >= total * (g + 1) || left == todo))) {
      ReplicaGroup G;
      G.rep0 = r0;
      G.nrep = r + 1 - r0;
      const Replica& R0 = B->reps[r0];
      G.blk0 = R0.blk0;
      G.mblk0 = R0.mblk0;
      G.nb = G.mnb = 0;
      G.bytes = 0;
      for (int k = r0; k <= r; ++k) {
        G.nb += B->reps[k].nb;
        G.mnb += B->reps[k].mnb;
        G.bytes += by[k];
      }
      B->groups.push_back(G);
      r0 = r + 1;
    }
  }
}

// Block tables for launch shape tb: replica r owns blocks [blk0, blk0 + nb)
// of every update launch.  Partial-sum slots are sized for the smallest
// shape at creation (pstride), so a shape change needs no reallocation.
static int batch_layout(lceq_batch* B, int tb, int xcd) {
  std::vector<int> blk2rep;
  int nb_total = 0;
  for (int r = 0; r < B->nrep; ++r) {
    Replica& R = B->reps[r];
    R.nb = (R.g.n + tb - 1) / tb;
    R.blk0 = nb_total;
    for (int k = 0; k < R.nb; ++k) blk2rep.push_back(r);
    nb_total += R.nb;
  }
  BHIP(hipMemcpyAsync(B->d_blk2rep, blk2rep.data(), blk2rep.size() * sizeof(int),
                      hipMemcpyHostToDevice, B->stream));
  B->total_nb = nb_total;
  B->tb = tb;
  B->xcd = xcd;
  batch_groups(B);
  if (B->graph) {
    BHIP(hipGraphExecDestroy(B->graph));
    B->graph = nullptr;
  }
  return batch_upload_table(B);
}


extern "C" int lceq_batch_create(int nrep, const int* dims, lceq_batch** out) {
  if (!out || !dims || nrep < 1) return bfail(LCEQ_ERR_ARG, "bad batch arguments");
  *out = nullptr;
  lceq_batch* B = new lceq_batch();
  B->nrep = nrep;
  B->reps.resize(nrep);
  B->field_off.resize(nrep);
  int64_t pool = 0, parts = 0;
  int nb_total = 0;
  std::vector<int> blk2rep, mblk2rep;
  for (int r = 0; r < nrep; ++r) {
    int lx = dims[3 * r], ly = dims[3 * r + 1], lz = dims[3 * r + 2];
    if (const char* why = geom_reject(lx, ly, lz)) {
      delete B;
      return bfail(LCEQ_ERR_ARG, "replica %d (%d x %d x %d): %s", r, lx, ly, lz, why);
    }
    Replica& R = B->reps[r];
    memset(&R, 0, sizeof(R));
    R.g = geom_init(lx, ly, lz);
    R.nb = (R.g.n + TB_MIN - 1) / TB_MIN;
    // one grid covers every replica's blocks (grid x below 2^31)
    if ((int64_t)nb_total + R.nb > INT32_MAX) {
      delete B;
      return bfail(LCEQ_ERR_ARG, "replicas 0..%d need 2^31 or more blocks of %d sites", r, TB_MIN);
    }
    R.blk0 = nb_total;
    R.pstride = R.nb;
    R.beta = 9.0;
    B->field_off[r] = pool;
    pool += 30 * (int64_t)R.g.npad;  // 24 link + 6 Higgs components
    parts += (int64_t)LCEQ_NSLOT * LCEQ_NQ * R.pstride;
    for (int k = 0; k < R.nb; ++k) blk2rep.push_back(r);
    nb_total += R.nb;
    R.mnb = (R.g.n + mlp::MLP_TB - 1) / mlp::MLP_TB;
    R.mblk0 = (int)mblk2rep.size();
    for (int k = 0; k < R.mnb; ++k) mblk2rep.push_back(r);
    B->total_sites += 2 * (int64_t)R.g.n;
  }
  B->total_nb = nb_total;
  B->total_mnb = (int)mblk2rep.size();
  B->pool_doubles = pool;
  B->parts_doubles = parts;
  B->hist_cap = 1024;
  B->group_bytes = DEFAULT_GROUP_BYTES;
  int rc = LCEQ_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (rc) return;
    if (hipMalloc(p, bytes) != hipSuccess) rc = bfail(LCEQ_ERR_HIP, "hipMalloc(%zu) failed", bytes);
  };
  if (hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess)
    rc = bfail(LCEQ_ERR_HIP, "hipStreamCreate failed");
  alloc((void**)&B->pool, pool * sizeof(double));
  alloc((void**)&B->parts, parts * sizeof(double));
  alloc((void**)&B->obs, (size_t)nrep * LCEQ_OBS_N * sizeof(double));
  alloc((void**)&B->hist, (size_t)B->hist_cap * nrep * LCEQ_OBS_N * sizeof(double));
  alloc((void**)&B->d_reps, nrep * sizeof(Replica));
  alloc((void**)&B->d_blk2rep, blk2rep.size() * sizeof(int));
  alloc((void**)&B->d_mblk2rep, mblk2rep.size() * sizeof(int));
  alloc((void**)&B->d_ctr, 2 * sizeof(uint64_t));
  if (rc) {
    lceq_batch_destroy(B);
    return rc;
  }
  int64_t poff = 0;
  for (int r = 0; r < nrep; ++r) {
    Replica& R = B->reps[r];
    R.U = B->pool + B->field_off[r];
    R.A = R.U + 24 * (size_t)R.g.npad;
    R.part = B->parts + poff;
    poff += (int64_t)LCEQ_NSLOT * LCEQ_NQ * R.pstride;
    R.obs = B->obs + (size_t)r * LCEQ_OBS_N;
    R.key = Philox{(uint32_t)(12345 + r), 0u};
  }
  (void)hipMemsetAsync(B->pool, 0, pool * sizeof(double), B->stream);
  (void)hipMemsetAsync(B->parts, 0, parts * sizeof(double), B->stream);
  std::vector<double> obs0((size_t)nrep * LCEQ_OBS_N, 0.0);
  for (int r = 0; r < nrep; ++r) obs0[(size_t)r * LCEQ_OBS_N + OBS_DELTA] = 0.4;
  (void)hipMemcpyAsync(B->obs, obs0.data(), obs0.size() * sizeof(double), hipMemcpyHostToDevice,
                       B->stream);
  (void)hipMemsetAsync(B->d_ctr, 0, 2 * sizeof(uint64_t), B->stream);
  (void)hipMemcpyAsync(B->d_mblk2rep, mblk2rep.data(), mblk2rep.size() * sizeof(int),
                       hipMemcpyHostToDevice, B->stream);
  if (batch_layout(B, 256, 1) != LCEQ_OK) {
    lceq_batch_destroy(B);
    return bfail(LCEQ_ERR_HIP, "batch initialisation failed");
  }
  *out = B;
  return LCEQ_OK;
}

extern "C" int lceq_batch_destroy(lceq_batch* B) {
  if (!B) return LCEQ_OK;
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  if (B->graph) (void)hipGraphExecDestroy(B->graph);
  void* bufs[] = {B->pool, B->parts, B->obs, B->hist, B->d_reps, B->d_blk2rep, B->d_mblk2rep,
                  B->d_ctr};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return LCEQ_OK;
}

extern "C" int lceq_batch_info(lceq_batch* B, int* nrep, int64_t* total_sites) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  if (nrep) *nrep = B->nrep;
  if (total_sites) *total_sites = B->total_sites;
  return LCEQ_OK;
}

extern "C" int lceq_batch_set_couplings(lceq_batch* B, const double* coup) {
  if (!B || !coup) return bfail(LCEQ_ERR_ARG, "null argument");
  for (int r = 0; r < B->nrep; ++r) {
    B->reps[r].beta = coup[4 * r];
    B->reps[r].beta_a = coup[4 * r + 1];
    B->reps[r].beta_2 = coup[4 * r + 2];
    B->reps[r].beta_4 = coup[4 * r + 3];
  }
  return batch_upload_table(B);
}

extern "C" int lceq_batch_rng_philox(lceq_batch* B, const uint64_t* seeds) {
  if (!B || !seeds) return bfail(LCEQ_ERR_ARG, "null argument");
  for (int r = 0; r < B->nrep; ++r)
    B->reps[r].key = Philox{(uint32_t)seeds[r], (uint32_t)(seeds[r] >> 32)};
  uint64_t zero[2] = {0, 0};
  BHIP(hipMemcpyAsync(B->d_ctr, zero, sizeof(zero), hipMemcpyHostToDevice, B->stream));
  B->sweeps_done = 0;
  return batch_upload_table(B);
}

extern "C" int lceq_batch_hotstart(lceq_batch* B, const uint64_t* seeds, double sigma) {
  if (!B || !seeds) return bfail(LCEQ_ERR_ARG, "null argument");
  // one single-lattice hot start per replica, on a view of the pooled storage
  for (int r = 0; r < B->nrep; ++r) {
    lceq_lattice view;
    memset(&view, 0, sizeof(view));
    view.g = B->reps[r].g;
    view.U = B->reps[r].U;
    view.A = B->reps[r].A;
    view.stream = B->stream;
    launch_hotstart(&view, seeds[r], sigma);
  }
  BHIP(hipGetLastError());
  BHIP(hipStreamSynchronize(B->stream));
  return LCEQ_OK;
}

static int batch_field(lceq_batch* B, int r, double* host, const double* hin, int64_t ld,
                       int links) {
  if (!B || r < 0 || r >= B->nrep) return bfail(LCEQ_ERR_ARG, "bad replica index");
  const Replica& R = B->reps[r];
  if (ld < R.g.n) return bfail(LCEQ_ERR_ARG, "ld < N");
  int ncomp = links ? 4 : 3, nclass = links ? 6 : 2;
  double* dev = links ? R.U : R.A;
  std::vector<double> st((size_t)nclass * ncomp * R.g.npad, 0.0);
  if (hin) {
    for (int cls = 0; cls < nclass; ++cls)
      for (int s = 0; s < R.g.n; ++s)
        for (int c = 0; c < ncomp; ++c)
          st[field_offset(ncomp, cls, c, s, R.g.npad)] = hin[((size_t)cls * ld + s) * ncomp + c];
    BHIP(hipMemcpyAsync(dev, st.data(), st.size() * sizeof(double), hipMemcpyHostToDevice,
                        B->stream));
    BHIP(hipStreamSynchronize(B->stream));
  } else {
    BHIP(hipMemcpyAsync(st.data(), dev, st.size() * sizeof(double), hipMemcpyDeviceToHost,
                        B->stream));
    BHIP(hipStreamSynchronize(B->stream));
    for (int cls = 0; cls < nclass; ++cls)
      for (int s = 0; s < R.g.n; ++s)
        for (int c = 0; c < ncomp; ++c)
          host[((size_t)cls * ld + s) * ncomp + c] = st[field_offset(ncomp, cls, c, s, R.g.npad)];
  }
  return LCEQ_OK;
}
extern "C" int lceq_batch_put_links(lceq_batch* B, int r, const double* u, int64_t ld) {
  return batch_field(B, r, nullptr, u, ld, 1);
}
extern "C" int lceq_batch_get_links(lceq_batch* B, int r, double* u, int64_t ld) {
  return batch_field(B, r, u, nullptr, ld, 1);
}
extern "C" int lceq_batch_put_higgs(lceq_batch* B, int r, const double* a, int64_t ld) {
  return batch_field(B, r, nullptr, a, ld, 0);
}
extern "C" int lceq_batch_get_higgs(lceq_batch* B, int r, double* a, int64_t ld) {
  return batch_field(B, r, a, nullptr, ld, 0);
}

template <int TBX>
static void enqueue_links(lceq_batch* B, const BatchCommon& c, int n_or, int measflag) {
  dim3 grid(c.total_nb), blk(TBX);
  int pass = 0;
  for (int k = -1; k < n_or; ++k) {
    int kp = k < 0;
    int m = measflag && (kp ? n_or == 0 : k == n_or - 1);
    for (int eo = 0; eo < 2; ++eo)
      for (int mu = 0; mu < 3; ++mu) {
        if (kp)
          hipLaunchKernelGGL((k_link_batch<true, TBX>), grid, blk, 0, B->stream, c, eo, mu, pass,
                             m, m);
        else
          hipLaunchKernelGGL((k_link_batch<false, TBX>), grid, blk, 0, B->stream, c, eo, mu, pass,
                             m, m);
        ++pass;
      }
  }
}

// ahiggs1 of parity eo for the replicas of c at pass ctr[1] + pass_in_sweep
template <int TBX>
static void enqueue_higgs(lceq_batch* B, const BatchCommon& c, int eo, int pass_in_sweep,
                          int measflag) {
  hipLaunchKernelGGL((k_higgs_batch<TBX>), dim3(c.total_nb), dim3(TBX), 0, B->stream, c, eo,
                     pass_in_sweep, measflag);
}

// launches over every replica
static BatchCommon batch_common(const lceq_batch* B) {
  return BatchCommon{B->d_reps, B->d_blk2rep, B->total_nb, 0, B->d_ctr, B->xcd};
}
// launches over one replica group
static BatchCommon group_common(const lceq_batch* B, const ReplicaGroup& G) {
  return BatchCommon{B->d_reps, B->d_blk2rep, G.nb, G.blk0, B->d_ctr, B->xcd};
}

// Enqueue one sweep of every replica: for each replica group in turn, 1 KP
// + n_or overrelaxation link sweeps + the Higgs update (ahiggs1, or the
// learned proposal of `pack`) and the group's per-replica finalize; then the
// counter tick.  Passes as a single lattice's sweep: 6 per link sweep, then
// one per Higgs parity, the same for every group (the counters tick once).
static void enqueue_sweep(lceq_batch* B, int n_or, int measflag,
                          const mlp::MlpPack* pack = nullptr) {
  int hp = 6 * (1 + n_or);  // pass of the first Higgs parity
  int link_kind = measflag ? (n_or ? 2 : 1) : 0;
  for (const ReplicaGroup& G : B->groups) {
    BatchCommon c = group_common(B, G);
    switch (B->tb) {
      case 64: enqueue_links<64>(B, c, n_or, measflag); break;
      case 128: enqueue_links<128>(B, c, n_or, measflag); break;
      case 512: enqueue_links<512>(B, c, n_or, measflag); break;
      default: enqueue_links<256>(B, c, n_or, measflag); break;
    }
    for (int eo = 0; eo < 2; ++eo) {
      if (pack) {
        hipLaunchKernelGGL(k_higgs_mlp_batch, dim3(G.mnb), dim3(mlp::MLP_TB), 0, B->stream, c,
                           B->d_mblk2rep, G.mnb, G.mblk0, pack, eo, hp + eo);
        continue;
      }
      switch (B->tb) {
        case 64: enqueue_higgs<64>(B, c, eo, hp + eo, measflag); break;
        case 128: enqueue_higgs<128>(B, c, eo, hp + eo, measflag); break;
        case 512: enqueue_higgs<512>(B, c, eo, hp + eo, measflag); break;
        default: enqueue_higgs<256>(B, c, eo, hp + eo, measflag); break;
      }
    }
    hipLaunchKernelGGL(k_finalize_batch, dim3(G.nrep), dim3(FIN_TB), 0, B->stream, B->d_reps,
                       G.rep0, link_kind, measflag, pack ? 2 : 1, B->d_ctr, B->hist, B->hist_cap,
                       B->nrep);
  }
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, B->stream, B->d_ctr, 1, hp + 2);
}

// nsweep sweeps, GRAPH_SWEEPS at a time from a captured graph (cached per
// schedule: n_or, measflag and the Higgs update), the rest launched directly
static int batch_run(lceq_batch* B, int nsweep, int n_or, int measflag, int use_graph,
                     const mlp::MlpPack* pack) {
  int direct = nsweep;
  if (use_graph && nsweep >= GRAPH_SWEEPS) {
    if (!B->graph || B->graph_n_or != n_or || B->graph_meas != measflag ||
        B->graph_pack != pack) {
      if (B->graph) {
        BHIP(hipGraphExecDestroy(B->graph));
        B->graph = nullptr;
      }
      hipGraph_t g;
      BHIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
      for (int k = 0; k < GRAPH_SWEEPS; ++k) enqueue_sweep(B, n_or, measflag, pack);
      BHIP(hipStreamEndCapture(B->stream, &g));
      hipError_t e = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        B->graph = nullptr;
        return bfail(LCEQ_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      B->graph_n_or = n_or;
      B->graph_meas = measflag;
      B->graph_pack = pack;
    }
    for (int it = 0; it < nsweep / GRAPH_SWEEPS; ++it) BHIP(hipGraphLaunch(B->graph, B->stream));
    direct = nsweep % GRAPH_SWEEPS;
  }
  for (int it = 0; it < direct; ++it) enqueue_sweep(B, n_or, measflag, pack);
  BHIP(hipGetLastError());
  B->sweeps_done += nsweep;
  return LCEQ_OK;
}

extern "C" int lceq_batch_sweep(lceq_batch* B, int nsweep, int n_or, const double* delta,
                                int measflag, int use_graph) {
  if (!B || nsweep < 0 || n_or < 0) return bfail(LCEQ_ERR_ARG, "bad sweep arguments");
  if (delta) {
    for (int r = 0; r < B->nrep; ++r)
      BHIP(hipMemcpyAsync(B->obs + (size_t)r * LCEQ_OBS_N + OBS_DELTA, delta + r, sizeof(double),
                          hipMemcpyHostToDevice, B->stream));
    BHIP(hipStreamSynchronize(B->stream));
  }
  return batch_run(B, nsweep, n_or, measflag, use_graph, nullptr);
}

extern "C" int lceq_batch_sweep_mlp(lceq_batch* B, lceq_mlp* M, int nsweep, int n_or,
                                    int measflag, int use_graph) {
  if (!B || !M || nsweep < 0 || n_or < 0) return bfail(LCEQ_ERR_ARG, "bad sweep arguments");
  return batch_run(B, nsweep, n_or, measflag, use_graph, mlp::device_pack(M));
}

// Training records from every replica at its own couplings, as
// lceq_mlp_collect does for one lattice: per parity, the inputs at the
// current F, nsteps ahiggs1 steps of that parity for every replica (kept in
// the chains; each takes one Philox pass, no history row), the new a.
extern "C" int lceq_batch_mlp_collect(lceq_batch* B, lceq_mlp* M, int nsteps, int stride) {
  if (!B || !M) return bfail(LCEQ_ERR_ARG, "null argument");
  if (nsteps < 1 || stride < 1) return bfail(LCEQ_ERR_ARG, "nsteps and stride must be >= 1");
  std::vector<int64_t> off(B->nrep + 1, 0);
  for (int r = 0; r < B->nrep; ++r) off[r + 1] = off[r] + (B->reps[r].g.n + stride - 1) / stride;
  const int64_t nrec = off[B->nrep];
  float* base = mlp::reserve_records(M, 2 * nrec);
  if (!base) return LCEQ_ERR_HIP;
  BatchCommon c = batch_common(B);
  for (int eo = 0; eo < 2; ++eo) {
    for (int target = 0; target < 2; ++target) {
      for (int r = 0; r < B->nrep; ++r) {
        const Replica& R = B->reps[r];
        mlp::launch_record(B->stream, R.g, R.U, R.A, eo, stride, (int)(off[r + 1] - off[r]),
                           R.beta_a, R.beta_2, R.beta_4,
                           base + (size_t)(eo * nrec + off[r]) * mlp::NREC, target);
      }
      if (target) break;
      for (int h = 0; h < nsteps; ++h) {
        switch (B->tb) {
          case 64: enqueue_higgs<64>(B, c, eo, 0, 0); break;
          case 128: enqueue_higgs<128>(B, c, eo, 0, 0); break;
          case 512: enqueue_higgs<512>(B, c, eo, 0, 0); break;
          default: enqueue_higgs<256>(B, c, eo, 0, 0); break;
        }
        hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, B->stream, B->d_ctr, 0, 1);
      }
    }
  }
  BHIP(hipGetLastError());
  BHIP(hipStreamSynchronize(B->stream));
  mlp::commit_records(M, 2 * nrec);
  return LCEQ_OK;
}

extern "C" int lceq_batch_sync(lceq_batch* B) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  BHIP(hipStreamSynchronize(B->stream));
  return LCEQ_OK;
}

// rows [sweeps_done - n, sweeps_done) of the per-replica history, as
// out[(k*nrep + r)*16 + i]; n <= min(sweeps measured, 1024)
extern "C" int lceq_batch_history(lceq_batch* B, double* out, int n) {
  if (!B || !out || n < 0 || n > B->hist_cap || (uint64_t)n > B->sweeps_done)
    return bfail(LCEQ_ERR_ARG, "bad history request");
  BHIP(hipStreamSynchronize(B->stream));
  size_t row = (size_t)B->nrep * LCEQ_OBS_N;
  std::vector<double> h((size_t)B->hist_cap * row);
  BHIP(hipMemcpy(h.data(), B->hist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < n; ++k) {
    uint64_t sw = B->sweeps_done - n + k;
    memcpy(out + k * row, h.data() + (sw % B->hist_cap) * row, row * sizeof(double));
  }
  return LCEQ_OK;
}

extern "C" int lceq_batch_observables(lceq_batch* B, double* out) {
  if (!B || !out) return bfail(LCEQ_ERR_ARG, "null argument");
  BHIP(hipStreamSynchronize(B->stream));
  BHIP(hipMemcpy(out, B->obs, (size_t)B->nrep * LCEQ_OBS_N * sizeof(double),
                 hipMemcpyDeviceToHost));
  return LCEQ_OK;
}

extern "C" int lceq_batch_set_shape(lceq_batch* B, int threads_per_block, int xcd_order) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  int t = threads_per_block;
  if (t != 64 && t != 128 && t != 256 && t != 512)
    return bfail(LCEQ_ERR_ARG, "threads per block must be 64, 128, 256 or 512");
  BHIP(hipStreamSynchronize(B->stream));
  return batch_layout(B, t, xcd_order ? 1 : 0);
}

extern "C" int lceq_batch_get_shape(lceq_batch* B, int* threads_per_block, int* xcd_order) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  if (threads_per_block) *threads_per_block = B->tb;
  if (xcd_order) *xcd_order = B->xcd;
  return LCEQ_OK;
}

extern "C" int lceq_batch_set_group_bytes(lceq_batch* B, int64_t group_bytes) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  if (group_bytes < 0) return bfail(LCEQ_ERR_ARG, "group_bytes must be >= 0 (0: one group)");
  BHIP(hipStreamSynchronize(B->stream));
  B->group_bytes = group_bytes;
  return batch_layout(B, B->tb, B->xcd);
}

extern "C" int lceq_batch_get_groups(lceq_batch* B, int64_t* group_bytes, int* ngroups,
                                     int* first_replica, int max_groups) {
  if (!B) return bfail(LCEQ_ERR_ARG, "null batch");
  if (group_bytes) *group_bytes = B->group_bytes;
  if (ngroups) *ngroups = (int)B->groups.size();
  if (first_replica)
    for (int g = 0; g < (int)B->groups.size() && g < max_groups; ++g)
      first_replica[g] = B->groups[g].rep0;
  return LCEQ_OK;
}

// Time every launch shape on `trials` sweeps of this batch (after one
// untimed sweep per shape) and keep the fastest.  The fields, observables
// and sweep counter are restored afterwards, so tuning leaves the chains
// where they were.  ms (optional, 8 values) receives the time per sweep of
// shapes (64, 128, 256, 512) x (launch order, XCD order).
extern "C" int lceq_batch_autotune(lceq_batch* B, int n_or, int trials, double* ms) {
  if (!B || trials < 1 || n_or < 0) return bfail(LCEQ_ERR_ARG, "bad autotune arguments");
  BHIP(hipStreamSynchronize(B->stream));
  double *pool0 = nullptr, *obs0 = nullptr;
  uint64_t* sw0 = nullptr;
  size_t pbytes = B->pool_doubles * sizeof(double), obytes = (size_t)B->nrep * LCEQ_OBS_N * sizeof(double);
  BHIP(hipMalloc((void**)&pool0, pbytes));
  int rc = LCEQ_OK;
  if (hipMalloc((void**)&obs0, obytes) != hipSuccess || hipMalloc((void**)&sw0, 16) != hipSuccess)
    rc = bfail(LCEQ_ERR_HIP, "hipMalloc failed");
  if (!rc && (hipMemcpyAsync(pool0, B->pool, pbytes, hipMemcpyDeviceToDevice, B->stream) ||
              hipMemcpyAsync(obs0, B->obs, obytes, hipMemcpyDeviceToDevice, B->stream) ||
              hipMemcpyAsync(sw0, B->d_ctr, 16, hipMemcpyDeviceToDevice, B->stream)))
    rc = bfail(LCEQ_ERR_HIP, "snapshot failed");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) || hipEventCreate(&e1))) rc = bfail(LCEQ_ERR_HIP, "events");
  const int shapes[4] = {64, 128, 256, 512};
  double best = 1e300;
  int best_tb = 256, best_x = 1;
  for (int i = 0; i < 8 && !rc; ++i) {
    int tb = shapes[i / 2], x = i % 2;
    if ((rc = batch_layout(B, tb, x))) break;
    enqueue_sweep(B, n_or, 0);
    (void)hipEventRecord(e0, B->stream);
    for (int t = 0; t < trials; ++t) enqueue_sweep(B, n_or, 0);
    (void)hipEventRecord(e1, B->stream);
    float el = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&el, e0, e1) != hipSuccess) {
      rc = bfail(LCEQ_ERR_HIP, "timing failed");
      break;
    }
    double per = (double)el / trials;
    if (ms) ms[i] = per;
    if (per < best) {
      best = per;
      best_tb = tb;
      best_x = x;
    }
  }
  if (!rc && (hipMemcpyAsync(B->pool, pool0, pbytes, hipMemcpyDeviceToDevice, B->stream) ||
              hipMemcpyAsync(B->obs, obs0, obytes, hipMemcpyDeviceToDevice, B->stream) ||
              hipMemcpyAsync(B->d_ctr, sw0, 16, hipMemcpyDeviceToDevice, B->stream) ||
              hipStreamSynchronize(B->stream)))
    rc = bfail(LCEQ_ERR_HIP, "restore failed");
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(pool0);
  if (obs0) (void)hipFree(obs0);
  if (sw0) (void)hipFree(sw0);
  if (rc) return rc;
  return batch_layout(B, best_tb, best_x);
}

