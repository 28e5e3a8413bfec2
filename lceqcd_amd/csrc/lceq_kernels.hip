// lceq_kernels.hip -- gfx950 kernels of the lattice update and measurement
// path.  One thread owns one site of the parity being updated; a (parity, mu)
// link class or a Higgs parity is one launch, so the checkerboard replaces the
// reference's MPI halo exchanges and barriers.  All arithmetic is FP64.
#include <math.h>

#include <algorithm>

#include "lceq_internal.hpp"
#include "lceq_sites.hpp"

namespace lceq {

static constexpr int TB = 256;

int blocks_for(int n) { return (n + TB - 1) / TB; }

// deterministic sum of x[0..n) by one 256-thread block; result in all threads
__device__ double block_sum_range(const double* __restrict__ x, int n) {
  __shared__ double red[4];
  double t = 0.0;
  for (int i = threadIdx.x; i < n; i += TB) t += x[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
  __syncthreads();
  double r = ((red[0] + red[1]) + red[2]) + red[3];
  __syncthreads();
  return r;
}

// ================================================================ links

struct LinkArgs {
  LinkSite site;
  int reduce;
  double* part;  // partial sums of this launch's slot: [LCEQ_NQ][pstride]
  int pstride, nb, xcd;
  SiteOrder order;
  int early;  // the schedule of link_update_site's EARLY (early_schedule)
};

// Schedule of the update launches (link_update_site, higgs_update_site):
// the early issue and the fenced uniform phase measured faster from 48^3 to
// 128^3 (4-6% per sweep), where a launch has 512 waves or more (32768 sites
// per class) and the lattice's link and Higgs fields fit the 256 MB Infinity
// Cache; smaller and larger lattices keep the previous schedule
static bool early_schedule(const lceq_lattice* L) {
  return L->g.n >= 32768 && 30.0 * 8.0 * (double)L->g.npad <= 256.0 * 1024.0 * 1024.0;
}

// Block index in the launch's block order: XCD-aware (contiguous sites per
// XCD) or the launch order
__device__ __forceinline__ int shape_block(int nb, int xcd) {
  int b = xcd_block(nb);  // computed either way: no branch before the loads
  return xcd ? b : (int)blockIdx.x;
}

// One site per thread, TBX sites per block
template <bool KP, bool FED, int TBX, int MU, bool EARLY>
__global__ __launch_bounds__(TBX) void k_link(LinkArgs a) {
  int blk = shape_block(a.nb, a.xcd);
  unsigned P = (unsigned)(blk * TBX + threadIdx.x);
  double red[3] = {0.0, 0.0, 0.0};
  if (P < (unsigned)a.site.g.n) {
    LinkSite p = a.site;
    p.mu = MU;
    link_update_site<KP, FED, EARLY>(p, order_site(p.g, a.order, P), red);
  }
  if (a.reduce) block_sum_store<3, TBX / 64>(red, a.part, a.pstride, blk);
}

int shape_blocks(const lceq_lattice* L) { return (L->g.n + L->tb - 1) / L->tb; }

template <bool KP, bool FED, int MU, bool EARLY>
static void link_dispatch_tb(lceq_lattice* L, const LinkArgs& a) {
  dim3 grid(a.nb);
  switch (L->tb) {
    case 64: hipLaunchKernelGGL((k_link<KP, FED, 64, MU, EARLY>), grid, dim3(64), 0, L->stream, a); break;
    case 128: hipLaunchKernelGGL((k_link<KP, FED, 128, MU, EARLY>), grid, dim3(128), 0, L->stream, a); break;
    case 512: hipLaunchKernelGGL((k_link<KP, FED, 512, MU, EARLY>), grid, dim3(512), 0, L->stream, a); break;
    default: hipLaunchKernelGGL((k_link<KP, FED, 256, MU, EARLY>), grid, dim3(256), 0, L->stream, a); break;
  }
}

template <bool KP, bool FED, int MU>
static void link_dispatch_mu(lceq_lattice* L, const LinkArgs& a) {
  if (a.early)
    link_dispatch_tb<KP, FED, MU, true>(L, a);
  else
    link_dispatch_tb<KP, FED, MU, false>(L, a);
}

// The direction is a template argument: the neighbour steps and link offsets
// of the staple are then fixed at compile time.
template <bool KP, bool FED>
static void link_dispatch(lceq_lattice* L, const LinkArgs& a) {
  switch (a.site.mu) {
    case 0: link_dispatch_mu<KP, FED, 0>(L, a); break;
    case 1: link_dispatch_mu<KP, FED, 1>(L, a); break;
    default: link_dispatch_mu<KP, FED, 2>(L, a); break;
  }
}

void launch_link(lceq_lattice* L, int kp, int eo, int mu, int measflag, int renorm_on_write,
                 const double* fed, uint64_t pass, int reduce) {
  LinkArgs a;
  a.site.g = L->g;
  a.site.U = L->U;
  a.site.A = L->A;
  a.site.eo = eo;
  a.site.mu = mu;
  a.site.beta = L->beta;
  a.site.beta_a = L->beta_a;
  a.site.fed = fed;
  a.site.key = Philox{(uint32_t)L->seed, (uint32_t)(L->seed >> 32)};
  a.site.pass = pass;
  a.site.measflag = measflag;
  a.site.renorm = renorm_on_write;
  a.reduce = reduce;
  a.part = L->partials + (size_t)(eo * 3 + mu) * LCEQ_NQ * L->pstride;
  a.pstride = L->pstride;
  a.nb = shape_blocks(L);
  a.xcd = L->xcd;
  a.order = site_order(L->g, L->strip);
  a.early = early_schedule(L) ? 1 : 0;
  int kc = kp ? KC_LINK_KP : KC_LINK_OR;
  prof_begin(L, kc);
  prof_count(L, kc);
  if (kp) {
    if (fed)
      link_dispatch<true, true>(L, a);
    else
      link_dispatch<true, false>(L, a);
  } else {
    if (fed)
      link_dispatch<false, true>(L, a);
    else
      link_dispatch<false, false>(L, a);
  }
  prof_end(L, kc);
}

// ============================================================ finalize

struct FinalizeArgs {
  const double* part;  // [LCEQ_NSLOT][LCEQ_NQ][pstride]
  int pstride, nb_link, nb_higgs;  // block partials per link / Higgs slot
  int link_kind, link_meas, higgs_kind, higgs_meas;
  double vol;
  double* obs;
  double* hist_row;
};

// One 1024-thread block sums every partial the closing routines wrote: six
// link slots x 3 quantities and two Higgs slots x 4 quantities.  Wave w owns
// quantities w and w+16; each lane sums its strided share of the block
// partials (the loads of both quantities in flight at once), a xor-butterfly
// finishes the wave sum (fixed order, so the result is reproducible), and
// thread 0 combines the slots in the reference's loop order: f/agauge.f:195-200
// (links), f/ahiggs1.f:107-123 and f/ahiggs2.f:138-143 (Higgs).  The
// observables are read by wave 0 alongside the partials, updated in LDS and
// written back (and to the history row) by 16 lanes at once.
constexpr int FIN_TB = 1024;
__global__ __launch_bounds__(1024) void k_finalize(FinalizeArgs p) {
  constexpr int NV = 6 * 3 + 2 * 4;
  __shared__ double tot[32];
  __shared__ double ob[LCEQ_OBS_N];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double ov = threadIdx.x < LCEQ_OBS_N ? p.obs[threadIdx.x] : 0.0;
  const double* x[2];
  int nbq[2];
  bool on[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int q = w + 16 * h;  // 0..17 link (slot*3 + quantity), 18..25 Higgs
    on[h] = q < NV && (q < 18 ? p.link_kind != 0 : p.higgs_kind != 0);
    size_t row = q < 18 ? (size_t)(q / 3) * LCEQ_NQ + q % 3
                        : (size_t)(6 + (q - 18) / 4) * LCEQ_NQ + (q - 18) % 4;
    x[h] = p.part + row * p.pstride;
    nbq[h] = on[h] ? (q < 18 ? p.nb_link : p.nb_higgs) : 0;
  }
  double t[2] = {0.0, 0.0};
  for (int i0 = lane; i0 < max(nbq[0], nbq[1]); i0 += 64 * 8) {
    double y[2][8];  // sixteen loads in flight before the adds
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int u = 0; u < 8; ++u) y[h][u] = i0 + 64 * u < nbq[h] ? x[h][i0 + 64 * u] : 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int u = 0; u < 8; ++u) t[h] += y[h][u];
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t[h] += __shfl_xor(t[h], off, 64);
    if (lane == 0 && w + 16 * h < NV) tot[w + 16 * h] = t[h];
  }
  if (threadIdx.x < LCEQ_OBS_N) ob[threadIdx.x] = ov;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double* v = tot;
    double* obs = ob;
    if (p.link_kind) {
      double plaq = 0.0, nacch = 0.0, acc1 = 0.0;
      for (int slot = 0; slot < 6; ++slot) {
        plaq = plaq + v[slot * 3];
        nacch += v[slot * 3 + 1];
        acc1 += v[slot * 3 + 2];
      }
      obs[OBS_NACCH] = nacch;
      obs[OBS_NACC1] = acc1;
      if (p.link_meas == 1) obs[OBS_PLAQ0] = plaq / (12.0 * p.vol);
      double accept;
      if (p.link_kind == 1) {  // (nacch/3V)(acc1/12V) with measflag, else nacch
        accept = nacch;
        if (p.link_meas == 1) accept = (accept / (3.0 * p.vol)) * (acc1 / (12.0 * p.vol));
      } else {
        accept = acc1 / (3.0 * p.vol);
      }
      obs[OBS_ACC_LINK] = accept;
    }
    if (p.higgs_kind) {
      double acc = v[18] + v[22];
      double r2 = v[19] + v[23], r4 = v[20] + v[24], kin = v[21] + v[25];
      if (p.higgs_meas == 1) {
        obs[OBS_A02] = r2 / p.vol;
        obs[OBS_A04] = r4 / p.vol;
        obs[OBS_A0KIN] = kin / (6.0 * p.vol);
      }
      double accept = acc / p.vol;
      if (p.higgs_kind == 1) {
        obs[OBS_ACC_HIGGS] = accept;
        double d = obs[OBS_DELTA];
        if (fabs(accept - 0.5) > 0.1) d = (accept > 0.5) ? 1.1131241 * d : 0.8989669 * d;
        obs[OBS_DELTA] = d;
      } else if (p.higgs_meas == 1) {
        obs[OBS_ACC_HIGGS] = accept;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < LCEQ_OBS_N) {
    double o = ob[threadIdx.x];
    p.obs[threadIdx.x] = o;
    if (p.hist_row) p.hist_row[threadIdx.x] = o;
  }
}

void launch_finalize(lceq_lattice* L, int link_kind, int link_meas, int higgs_kind,
                     int higgs_meas, double* hist_row, int higgs_nb) {
  FinalizeArgs a;
  a.part = L->partials;
  a.pstride = L->pstride;
  a.nb_link = shape_blocks(L);
  a.nb_higgs = higgs_nb > 0 ? higgs_nb : shape_blocks(L);
  a.link_kind = link_kind;
  a.link_meas = link_meas;
  a.higgs_kind = higgs_kind;
  a.higgs_meas = higgs_meas;
  a.vol = (double)L->vol;
  a.obs = L->obs;
  a.hist_row = hist_row;
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_TB), 0, L->stream, a);
}

// staple only, for the staples() entry point (f/staples.f:9)
__global__ __launch_bounds__(256) void k_staples(Geom g, const double* __restrict__ U,
                                                 double* __restrict__ Vst, int eo, int mu) {
  int s = blockIdx.x * TB + threadIdx.x;
  if (s >= g.n) return;
  Coord c = site_coord(g, eo, s);
  st_link(Vst, g, mu, eo, s, staple(U, g, eo, mu, s, c));
}

void launch_staples(lceq_lattice* L, int eo, int mu) {
  hipLaunchKernelGGL(k_staples, dim3(blocks_for(L->g.n)), dim3(TB), 0, L->stream, L->g, L->U,
                     L->Vst, eo, mu);
}

// ================================================================ Higgs

__global__ __launch_bounds__(256) void k_force(Geom g, const double* __restrict__ U,
                                               const double* __restrict__ A,
                                               double* __restrict__ F, int eo) {
  int s = blockIdx.x * TB + threadIdx.x;
  if (s >= g.n) return;
  st_higgs(F, g, 0, s, hop_force(U, A, g, eo, s, site_coord(g, eo, s)));
}

void launch_force(lceq_lattice* L, int eo) {
  hipLaunchKernelGGL(k_force, dim3(blocks_for(L->g.n)), dim3(TB), 0, L->stream, L->g, L->U, L->A,
                     L->F, eo);
}

struct HiggsArgs {
  HiggsSite site;
  double* part;
  int pstride, nb, xcd;
  SiteOrder order;
  int early;  // as LinkArgs::early
};

template <int KIND, bool FED, int TBX, bool EARLY>
__global__ __launch_bounds__(TBX) void k_higgs(HiggsArgs a) {
  int blk = shape_block(a.nb, a.xcd);
  unsigned P = (unsigned)(blk * TBX + threadIdx.x);
  double red[4] = {0.0, 0.0, 0.0, 0.0};
  if (P < (unsigned)a.site.g.n)
    higgs_update_site<KIND, FED, EARLY>(a.site, order_site(a.site.g, a.order, P), red);
  block_sum_store<4, TBX / 64>(red, a.part, a.pstride, blk);
}

template <int KIND, bool FED, bool EARLY>
static void higgs_dispatch_tb(lceq_lattice* L, const HiggsArgs& a) {
  dim3 grid(a.nb);
  switch (L->tb) {
    case 64: hipLaunchKernelGGL((k_higgs<KIND, FED, 64, EARLY>), grid, dim3(64), 0, L->stream, a); break;
    case 128: hipLaunchKernelGGL((k_higgs<KIND, FED, 128, EARLY>), grid, dim3(128), 0, L->stream, a); break;
    case 512: hipLaunchKernelGGL((k_higgs<KIND, FED, 512, EARLY>), grid, dim3(512), 0, L->stream, a); break;
    default: hipLaunchKernelGGL((k_higgs<KIND, FED, 256, EARLY>), grid, dim3(256), 0, L->stream, a); break;
  }
}

// the early issue exists for ahiggs1 only (EARLY = KIND == 1 there)
template <int KIND, bool FED>
static void higgs_dispatch(lceq_lattice* L, const HiggsArgs& a) {
  if (a.early)
    higgs_dispatch_tb<KIND, FED, KIND == 1>(L, a);
  else
    higgs_dispatch_tb<KIND, FED, false>(L, a);
}

void launch_higgs(lceq_lattice* L, int kind, int eo, int measflag, const double* fed,
                  uint64_t pass) {
  HiggsArgs a;
  a.site.g = L->g;
  a.site.U = L->U;
  a.site.A = L->A;
  a.site.eo = eo;
  a.site.imode = L->imode;
  a.site.beta_a = L->beta_a;
  a.site.beta_2 = L->beta_2;
  a.site.beta_4 = L->beta_4;
  a.site.delta = L->obs + OBS_DELTA;
  a.site.fed = fed;
  a.site.key = Philox{(uint32_t)L->seed, (uint32_t)(L->seed >> 32)};
  a.site.pass = pass;
  a.site.measflag = measflag;
  a.part = L->partials + (size_t)(6 + eo) * LCEQ_NQ * L->pstride;
  a.pstride = L->pstride;
  a.nb = shape_blocks(L);
  a.xcd = L->xcd;
  a.order = site_order(L->g, L->strip);
  a.early = early_schedule(L) ? 1 : 0;
  prof_begin(L, KC_HIGGS);
  prof_count(L, KC_HIGGS);
  if (kind == 1) {
    if (fed) higgs_dispatch<1, true>(L, a); else higgs_dispatch<1, false>(L, a);
  } else if (kind == 2) {
    if (fed) higgs_dispatch<2, true>(L, a); else higgs_dispatch<2, false>(L, a);
  } else {
    if (fed) higgs_dispatch<3, true>(L, a); else higgs_dispatch<3, false>(L, a);
  }
  prof_end(L, KC_HIGGS);
}

// =============================================================== renorm

// f/renorm.f:24-34 / f/u_f_renorm.f:93-103 over all 3V links
__global__ __launch_bounds__(256) void k_renorm(Geom g, double* __restrict__ U) {
  int s = blockIdx.x * TB + threadIdx.x;
  int cls = blockIdx.y;  // mu*2 + p
  if (s >= g.n) return;
  int mu = cls >> 1, p = cls & 1;
  st_link(U, g, mu, p, s, normalized(ld_link(U, g, mu, p, s)));
}

void launch_renorm(lceq_lattice* L, double* U) {
  hipLaunchKernelGGL(k_renorm, dim3(blocks_for(L->g.n), 6), dim3(TB), 0, L->stream, L->g, U);
}

// Field guard: per block of sites (both parities), the largest unitarity
// deviation | |alpha|^2 + |beta|^2 - 1 | of the finite links and the number
// of non-finite values among the links and the Higgs field:
// part[2 blk] = deviation, part[2 blk + 1] = count.
__global__ __launch_bounds__(256) void k_check_fields(Geom g, const double* __restrict__ U,
                                                      const double* __restrict__ A,
                                                      double* __restrict__ part) {
  __shared__ double wmax[4], wbad[4];
  int s = blockIdx.x * TB + threadIdx.x;
  double dev = 0.0, bad = 0.0;
  if (s < g.n) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int mu = 0; mu < 3; ++mu) {
        su2 u = ld_link(U, g, mu, p, s);
        double c[4] = {u.a.re, u.a.im, u.b.re, u.b.im};
#pragma unroll
        for (int k = 0; k < 4; ++k) bad += isfinite(c[k]) ? 0.0 : 1.0;
        double d = fabs(((c[0] * c[0] + c[1] * c[1]) + (c[2] * c[2] + c[3] * c[3])) - 1.0);
        if (isfinite(d)) dev = fmax(dev, d);
      }
      vec3 a = ld_higgs(A, g, p, s);
#pragma unroll
      for (int k = 0; k < 3; ++k) bad += isfinite(a.x[k]) ? 0.0 : 1.0;
    }
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
    part[2 * blockIdx.x] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
    part[2 * blockIdx.x + 1] = ((wbad[0] + wbad[1]) + wbad[2]) + wbad[3];
  }
}

void launch_check_fields(lceq_lattice* L, double* part) {
  hipLaunchKernelGGL(k_check_fields, dim3(blocks_for(L->g.n)), dim3(TB), 0, L->stream, L->g,
                     L->U, L->A, part);
}

// ========================================================= gauge fixing

struct GfArgs {
  Geom g;
  double* Uf;
  double* Af;
  int eo, compute_da, renorm, guarded;
  double lambda, alpha;
  double* obs;  // OBS_GF_DONE read when guarded; the convergence test writes
  double* part;
  int pstride, nb;
  // convergence test of the previous iteration at the start of this launch
  // (landau_gauge iterations after the first of a batch): `iter` iterations
  // are complete, their last da is in `part`
  int check;
  double iter, limit_or, max_iter;
};

// da1 = (sum of the per-block partials of the eo = 1 half) / N
// (f/or_gauge_fixing1.f:156-157), summed in a fixed order that does not
// depend on the caller; every thread of a 256-thread block must call it.
__device__ __forceinline__ double gfix_da(const double* __restrict__ part, int nb, double n) {
  double v[1] = {0.0};
  for (int i0 = threadIdx.x; i0 < nb; i0 += TB * 4) {
    double y[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) y[u] = i0 + TB * u < nb ? part[i0 + TB * u] : 0.0;
#pragma unroll
    for (int u = 0; u < 4; ++u) v[0] += y[u];
  }
  block_sum<1>(v);
  return v[0] / n;
}

// The loop condition of f/landau_gauge.f:54-59 after `iter` iterations
// whose last da is `da`, plus the iteration cap: returns whether to stop;
// one thread records da, the count and the flags.
__device__ __forceinline__ bool gfix_stop(double da, double iter, double limit_or,
                                          double max_iter, double* __restrict__ obs, bool writer) {
  bool conv = !(da > limit_or);
  bool stop = conv || iter >= max_iter;
  if (writer) {
    obs[OBS_DA] = da;
    obs[OBS_ITER] = iter;
    if (stop) obs[OBS_GF_DONE] = 1.0;
    if (stop && !conv) obs[OBS_GF_CAPPED] = 1.0;
  }
  return stop;
}

// One parity half of f/or_gauge_fixing1.f:54-328.  Thread x reads and
// writes only U_mu(x) and U_mu(x-mu) for mu = 1..3 and a_f(x), so the half
// step has no cross-thread hazards.  With renorm=1 (eo = 1 of a
// landau_gauge iteration) every link of the lattice is written exactly once
// and f/u_f_renorm.f is applied on that write.  With check=1 every block
// first repeats the convergence test of the previous iteration (the same
// sum in the same order, so all blocks agree) and the launch returns when
// the loop has ended; the operand loads are issued before the test.
__global__ __launch_bounds__(256) void k_gfix(GfArgs p) {
  const Geom& g = p.g;
  int blk = xcd_block(p.nb);
  int s = blk * TB + threadIdx.x;
  bool act = s < g.n;
  int eo = p.eo, oe = eo ^ 1;
  su2 ux[3], uy[3];
  int sy[3];
  vec3 af;
  if (act) {
    Steps st = site_steps(g, site_coord(g, eo, s));
#pragma unroll
    for (int mu = 0; mu < 3; ++mu) {
      sy[mu] = s + st.m[mu];
      ux[mu] = ld_link(p.Uf, g, mu, eo, s);
      uy[mu] = ld_link(p.Uf, g, mu, oe, sy[mu]);
    }
    af = ld_higgs(p.Af, g, eo, s);
  }
  // uniform over the grid; tested after the operand loads are issued, so that
  // its round trip overlaps theirs
  if (p.guarded && p.obs[OBS_GF_DONE] != 0.0) return;
  if (p.check) {
    double da = gfix_da(p.part, p.nb, (double)g.n);
    if (gfix_stop(da, p.iter, p.limit_or, p.max_iter, p.obs, blk == 0 && threadIdx.x == 0))
      return;  // block-uniform
  }
  double red[1] = {0.0};
  if (act) {
    // :97-107, :116-142
    su2 s1{c_add(ux[0].a, ux[1].a), c_add(ux[0].b, ux[1].b)};
    s1.a = c_add(s1.a, c_scale(p.alpha, ux[2].a));
    s1.b = c_add(s1.b, c_scale(p.alpha, ux[2].b));
    su2 s2{c_add(uy[0].a, uy[1].a), c_add(uy[0].b, uy[1].b)};
    s2.a = c_add(s2.a, c_scale(p.alpha, uy[2].a));
    s2.b = c_add(s2.b, c_scale(p.alpha, uy[2].b));
    su2 w{c_add(s1.a, c_conj(s2.a)), c_sub(s1.b, s2.b)};  // :145-146
    if (p.compute_da) {                                    // :153-158
      double d = c_sub(s1.a, s2.a).im;
      red[0] = d * d + (w.b.re * w.b.re + w.b.im * w.b.im);
    }
    double k = (w.a.re * w.a.re + w.a.im * w.a.im) + (w.b.re * w.b.re + w.b.im * w.b.im);
    double rk = rsqrt(k);  // w / sqrt(k), :169-183
    su2 o{c_scale(rk, w.a), c_scale(rk, w.b)};
    double ra = fmin(1.0, fmax(-1.0, o.a.re));
    double th = acos(ra);
    if (th != 0.0) {
      double tho = p.lambda * th;
      // sin(th) for th = acos(ra) in (0, pi] is sqrt((1 - ra)(1 + ra)); the
      // factored form keeps full relative precision as th -> 0
      double f = sin(tho) * rsqrt((1.0 - ra) * (1.0 + ra));
      double t2 = f * o.a.im;
      o.a = c_mk(cos(tho), -t2);
      o.b = c_mk(-(f * o.b.re), -(f * o.b.im));
    }
    st_higgs(p.Af, g, eo, s, adj_rot<false>(o, af));  // g X(a_f) g^+, :243-252
#pragma unroll
    for (int mu = 0; mu < 3; ++mu) {  // :256-324
      su2 nx = m_nn(o, ux[mu]);
      su2 ny = m_nd(uy[mu], o);
      if (p.renorm) {
        nx = normalized(nx);
        ny = normalized(ny);
      }
      st_link(p.Uf, g, mu, eo, s, nx);
      st_link(p.Uf, g, mu, oe, sy[mu], ny);
    }
  }
  if (p.compute_da) block_sum_store<1>(red, p.part, p.pstride, blk);
}

// da1 after one or_gauge_fixing call (guarded = 0), or the convergence test
// of a landau_gauge batch's last iteration (guarded = 1, `iter` iterations
// complete): the loop ends when da1 <= limit_or or at the iteration cap.
__global__ __launch_bounds__(256) void k_gfix_finalize(const double* __restrict__ part, int nb,
                                                       double n, int guarded, double iter,
                                                       double limit_or, double max_iter,
                                                       double* __restrict__ obs) {
  if (guarded && obs[OBS_GF_DONE] != 0.0) return;
  double da = gfix_da(part, nb, n);
  if (guarded)
    gfix_stop(da, iter, limit_or, max_iter, obs, threadIdx.x == 0);
  else if (threadIdx.x == 0)
    obs[OBS_DA] = da;
}

void launch_gfix(lceq_lattice* L, int eo, int compute_da, int renorm, int guarded, int check,
                 int iter) {
  GfArgs a;
  a.g = L->g;
  a.Uf = L->Uf;
  a.Af = L->Af;
  a.eo = eo;
  a.compute_da = compute_da;
  a.renorm = renorm;
  a.guarded = guarded;
  a.lambda = L->lambda;
  a.alpha = L->alpha;
  a.obs = L->obs;
  a.part = L->partials;
  a.pstride = L->pstride;
  a.nb = blocks_for(L->g.n);
  a.check = check;
  a.iter = (double)iter;
  a.limit_or = L->limit_or;
  a.max_iter = (double)L->max_iter;
  prof_begin(L, KC_GFIX);
  prof_count(L, KC_GFIX);
  hipLaunchKernelGGL(k_gfix, dim3(a.nb), dim3(TB), 0, L->stream, a);
  prof_end(L, KC_GFIX);
}

void launch_gfix_finalize(lceq_lattice* L, int guarded, int iter) {
  hipLaunchKernelGGL(k_gfix_finalize, dim3(1), dim3(TB), 0, L->stream, L->partials,
                     blocks_for(L->g.n), (double)L->g.n, guarded, (double)iter, L->limit_or,
                     (double)L->max_iter, L->obs);
}

// ========================================================== correlators

// z-slice sums over both parities (f/correlator.f:53-79, f/scalarcorr.f:43-56,
// f/polyakov.f:47-58), each slice split over `split` blocks (about 512
// sites each, corr_split).  Quantities q (partial j at
// slices[(q*lz + z)*split + j]):
//   0-2 a_f(1..3); 3+3*mu: Im alpha_mu, 4+3*mu: Im beta_mu, 5+3*mu: -Re beta_mu
//   (the (Im a, -i b) projection of u_f); 12 rho(a); 13 cos(4|a|).
__global__ __launch_bounds__(256) void k_slices(Geom g, int split,
                                                const double* __restrict__ Uf,
                                                const double* __restrict__ Af,
                                                const double* __restrict__ A,
                                                double* __restrict__ slices) {
  __shared__ double red[4][14];
  int z = blockIdx.x, j = blockIdx.y;
  int nz = 2 * g.h2, chunk = (nz + split - 1) / split;
  int i0 = j * chunk, i1 = min(nz, i0 + chunk);
  double acc[14];
#pragma unroll
  for (int q = 0; q < 14; ++q) acc[q] = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += TB) {
    int p = i >= g.h2, s = z * g.h2 + (i - p * g.h2);
    vec3 af = ld_higgs(Af, g, p, s), a = ld_higgs(A, g, p, s);
#pragma unroll
    for (int k = 0; k < 3; ++k) acc[k] += af.x[k];
#pragma unroll
    for (int mu = 0; mu < 3; ++mu) {
      su2 u = ld_link(Uf, g, mu, p, s);
      acc[3 + 3 * mu] += u.a.im;
      acc[4 + 3 * mu] += u.b.im;
      acc[5 + 3 * mu] -= u.b.re;
    }
    double rho = a.x[0] * a.x[0] + a.x[1] * a.x[1] + a.x[2] * a.x[2];
    acc[12] += rho;
    acc[13] += cos(4.0 * sqrt(rho));
  }
#pragma unroll
  for (int q = 0; q < 14; ++q) {
    double t = acc[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = t;
  }
  __syncthreads();
  if (threadIdx.x < 14) {
    int q = threadIdx.x;
    slices[(q * g.lz + z) * split + j] = ((red[0][q] + red[1][q]) + red[2][q]) + red[3][q];
  }
}

// Periodic z-correlators on the slice means (f/correlator.f:139-164,
// f/scalarcorr.f:96-104, f/polyakov.f:97-107).  corr layout:
// [0,lz) C_a, [lz,2lz) C_u, [2lz,3lz) C_3, [3lz,4lz) rho, [4lz,5lz) Polyakov,
// then mean(rho), mean(P)^2, a0f(1..3).  The slice partials are summed in a
// fixed order into slice means, held in LDS (dynamic, 14 lz doubles) when
// LDS is set, else in `slices` itself.
constexpr int CORR_TB = 1024;    // threads of the one k_corr block
constexpr int CORR_MAX_SPLIT = 16;

template <bool LDS>
__global__ __launch_bounds__(CORR_TB) void k_corr(Geom g, int split, double* __restrict__ slices,
                                                  double* __restrict__ corr) {
  extern __shared__ double lds_slices[];
  int lz = g.lz;
  double area = (double)(g.lx * g.ly);
  double* as = LDS ? lds_slices : slices;
  // in place (LDS = false) is safe: element i reads only its own partials,
  // and split = 1 then (see corr_split), so slices[i * split] = slices[i].
  // All partials of an element are loaded before they are added (in order).
  for (int i = threadIdx.x; i < 14 * lz; i += CORR_TB) {
    double y[CORR_MAX_SPLIT];
#pragma unroll
    for (int j = 0; j < CORR_MAX_SPLIT; ++j) y[j] = j < split ? slices[i * split + j] : 0.0;
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < CORR_MAX_SPLIT; ++j)
      if (j < split) v += y[j];
    as[i] = v / area;
  }
  __syncthreads();
  // one thread per (correlator, separation); each sum runs over x0 in order
  for (int task = threadIdx.x; task < 5 * lz; task += CORR_TB) {
    int kind = task / lz, z = task - kind * lz;
    double c = 0.0;
    // one loop per kind, unrolled so that the LDS reads of several x0 are in
    // flight together; the additions stay in x0 order
    if (kind == 0) {
#pragma unroll 8
      for (int x0 = 0; x0 < lz; ++x0) {
        int x1 = x0 + z < lz ? x0 + z : x0 + z - lz;
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) d = d + as[k * lz + x0] * as[k * lz + x1];
        c = c + d;
      }
    } else if (kind == 1) {  // Re[A1 A1' + A2 conj(A2')] for mu = 1,2, A1 real
#pragma unroll 8
      for (int x0 = 0; x0 < lz; ++x0) {
        int x1 = x0 + z < lz ? x0 + z : x0 + z - lz;
        double t = 0.0;
#pragma unroll
        for (int mu = 0; mu < 2; ++mu) {
          double a1 = as[(3 + 3 * mu) * lz + x0], a1p = as[(3 + 3 * mu) * lz + x1];
          double br = as[(4 + 3 * mu) * lz + x0], bi = as[(5 + 3 * mu) * lz + x0];
          double brp = as[(4 + 3 * mu) * lz + x1], bip = as[(5 + 3 * mu) * lz + x1];
          t = t + a1 * a1p;
          t = t + (br * brp + bi * bip);
        }
        c = c + t;
      }
    } else if (kind == 2) {  // mu = 3
#pragma unroll 8
      for (int x0 = 0; x0 < lz; ++x0) {
        int x1 = x0 + z < lz ? x0 + z : x0 + z - lz;
        double a1 = as[9 * lz + x0], a1p = as[9 * lz + x1];
        double br = as[10 * lz + x0], bi = as[11 * lz + x0];
        double brp = as[10 * lz + x1], bip = as[11 * lz + x1];
        c = c + (a1 * a1p + (br * brp + bi * bip));
      }
    } else {  // rho, Polyakov
      const double* q = as + (kind == 3 ? 12 : 13) * lz;
#pragma unroll 8
      for (int x0 = 0; x0 < lz; ++x0) {
        int x1 = x0 + z < lz ? x0 + z : x0 + z - lz;
        c = c + q[x0] * q[x1];
      }
    }
    corr[kind * lz + z] = kind == 1 ? c / (lz * 2) : c / lz;
  }
  // the means, one quantity per thread, summed over z in order
  if (threadIdx.x < 5) {
    const int qs[5] = {12, 13, 0, 1, 2};
    int q = qs[threadIdx.x];
    double m = 0.0;
#pragma unroll 8
    for (int z = 0; z < lz; ++z) m += as[q * lz + z];
    m = m / lz;
    // mean(rho), mean(P)^2, a0f(1..3) = mean over sites = mean of slice means
    corr[5 * lz + threadIdx.x] = threadIdx.x == 1 ? m * m : m;
  }
}

// slice means in LDS up to this many bytes (one block, 64 KB per workgroup)
constexpr size_t CORR_LDS_MAX = 64 * 1024;

int corr_split(const Geom& g) {
  if (14 * (size_t)g.lz * sizeof(double) > CORR_LDS_MAX) return 1;
  int nz = 2 * g.h2;
  // 512 sites per block, at most CORR_MAX_SPLIT blocks per slice.  One site
  // per thread (twice the blocks) was measured slower: 64^3 slice sums
  // 11.3 -> 13.4 us (twice the block reductions) and the correlator block
  // 12.5 -> 17.3 us (twice the partials per slice)
  return std::max(1, std::min(CORR_MAX_SPLIT, (nz + 2 * TB - 1) / (2 * TB)));
}

void launch_slices(lceq_lattice* L) {
  int split = corr_split(L->g);
  hipLaunchKernelGGL(k_slices, dim3(L->g.lz, split), dim3(TB), 0, L->stream, L->g, split, L->Uf,
                     L->Af, L->A, L->slices);
}
void launch_corr(lceq_lattice* L) {
  int split = corr_split(L->g);
  size_t lds = 14 * (size_t)L->g.lz * sizeof(double);
  if (lds <= CORR_LDS_MAX)
    hipLaunchKernelGGL(k_corr<true>, dim3(1), dim3(CORR_TB), lds, L->stream, L->g, split,
                       L->slices, L->corr);
  else
    hipLaunchKernelGGL(k_corr<false>, dim3(1), dim3(CORR_TB), 0, L->stream, L->g, split,
                       L->slices, L->corr);
}

// ============================================ array-level entry points

// f/heat_kpi.f:9-101 on n independent sites (b, a0_old in; W out as 4
// doubles per site in (alpha, beta) form); partial 0 of slot 0 counts accepts.
__global__ __launch_bounds__(256) void k_heat_kpi(int n, const double* __restrict__ b,
                                                  const double* __restrict__ a0_old,
                                                  double* __restrict__ v,
                                                  const double* __restrict__ fed, Philox key,
                                                  uint64_t pass, double* __restrict__ part,
                                                  int nb) {
  int s = blockIdx.x * TB + threadIdx.x;
  double red[1] = {0.0};
  if (s < n) {
    double r1, r2, r3, r4, r5, r6;
    if (fed) {
      r1 = fed[3 * s];
      r2 = fed[3 * s + 1];
      r3 = fed[3 * s + 2];
      r4 = fed[3 * n + s];
      r5 = fed[4 * n + s];
      r6 = fed[5 * n + s];
    } else {
      philox_pair(key, (uint32_t)s, 0, pass, 0, r1, r2);
      philox_pair(key, (uint32_t)s, 0, pass, 1, r3, r4);
      philox_pair(key, (uint32_t)s, 0, pass, 2, r5, r6);
    }
    if (r1 == 0.0) r1 = 0.0000033;
    if (r2 == 0.0) r2 = 0.0000066;
    double cc = cos(TPI * r3);
    double a0 = -(log(r1) + log(r2) * (cc * cc)) * b[s];
    bool lacc = r4 * r4 <= (1.0 - 0.5 * a0);
    red[0] = lacc ? 1.0 : 0.0;
    a0 = lacc ? 1.0 - a0 : a0_old[s];
    double a1 = 1.0 - a0 * a0;
    a1 = a1 >= 0.0 ? sqrt(a1) : 0.0;
    double a2 = 2.0 * (r5 - 0.5);
    double tt = 1.0 - a2 * a2;
    double a0b = tt >= 0.0 ? a1 * sqrt(tt) : 0.0;
    double ang = TPI * r6;
    v[4 * (size_t)s] = a0;
    v[4 * (size_t)s + 1] = a1 * a2;
    v[4 * (size_t)s + 2] = a0b * sin(ang);
    v[4 * (size_t)s + 3] = a0b * cos(ang);
  }
  block_sum_store<1>(red, part, nb, blockIdx.x);
}

void launch_heat_kpi(lceq_lattice* L, int n, const double* b, const double* a0_old, double* v,
                     const double* fed, uint64_t pass) {
  int nb = blocks_for(n);
  hipLaunchKernelGGL(k_heat_kpi, dim3(nb), dim3(TB), 0, L->stream, n, b, a0_old, v, fed,
                     Philox{(uint32_t)L->seed, (uint32_t)(L->seed >> 32)}, pass, L->partials, nb);
}

// f/gauss.f:10-31
__global__ __launch_bounds__(256) void k_gauss(int n, const double* __restrict__ r,
                                               double* __restrict__ x,
                                               const double* __restrict__ fed, Philox key,
                                               uint64_t pass) {
  int s = blockIdx.x * TB + threadIdx.x;
  if (s >= n) return;
  double w1, w2;
  if (fed) {
    w1 = fed[s];
    w2 = fed[n + s];
  } else {
    philox_pair(key, (uint32_t)s, 0, pass, 0, w1, w2);
  }
  double t = -1.0 * (1.0 / r[s]) * log(w1);
  x[s] = sqrt(t) * cos(TPI * w2);
}

void launch_gauss(lceq_lattice* L, int n, const double* r, double* x, const double* fed,
                  uint64_t pass) {
  hipLaunchKernelGGL(k_gauss, dim3(blocks_for(n)), dim3(TB), 0, L->stream, n, r, x, fed,
                     Philox{(uint32_t)L->seed, (uint32_t)(L->seed >> 32)}, pass);
}

// f/matmul.f MULTALL{NN,DN,ND,DD,AD,AN} on site vectors (kind 0..5); operands
// in (alpha, beta) quadruples, Higgs triples for AD/AN; s1/s2 may be null.
__global__ __launch_bounds__(256) void k_multall(int kind, const double* __restrict__ op1,
                                                 const double* __restrict__ op2,
                                                 double* __restrict__ out, int n,
                                                 const int* __restrict__ s1,
                                                 const int* __restrict__ s2) {
  int i = blockIdx.x * TB + threadIdx.x;
  if (i >= n) return;
  int i1 = s1 ? s1[i] : i, i2 = s2 ? s2[i] : i;
  const double* q2 = op2 + 4 * (size_t)i2;
  su2 q{c_mk(q2[0], q2[1]), c_mk(q2[2], q2[3])}, h;
  if (kind >= 4) {
    const double* a1 = op1 + 3 * (size_t)i1;
    vec3 a{{a1[0], a1[1], a1[2]}};
    h = kind == 4 ? m_ad(a, q) : m_an(a, q);
  } else {
    const double* q1 = op1 + 4 * (size_t)i1;
    su2 pp{c_mk(q1[0], q1[1]), c_mk(q1[2], q1[3])};
    h = kind == 0 ? m_nn(pp, q) : kind == 1 ? m_dn(pp, q) : kind == 2 ? m_nd(pp, q) : m_dd(pp, q);
  }
  double* o = out + 4 * (size_t)i;
  o[0] = h.a.re;
  o[1] = h.a.im;
  o[2] = h.b.re;
  o[3] = h.b.im;
}

void launch_multall(lceq_lattice* L, int kind, const double* op1, const double* op2, double* out,
                    int n, const int* s1, const int* s2) {
  hipLaunchKernelGGL(k_multall, dim3(blocks_for(n)), dim3(TB), 0, L->stream, kind, op1, op2, out,
                     n, s1, s2);
}

// ======================================================= start configs

// Haar links (four Box-Muller normals, stream 1) and N(0, sigma^2) Higgs
// components (stream 2); counter index = (mu*2+p)*n + s resp. p*n + s.
__global__ __launch_bounds__(256) void k_hotstart(Geom g, double* __restrict__ U,
                                                  double* __restrict__ A, Philox key,
                                                  double sigma) {
  int s = blockIdx.x * TB + threadIdx.x;
  int cls = blockIdx.y;  // 0..5 links, 6..7 Higgs
  if (s >= g.n) return;
  double u0, u1, u2, u3;
  if (cls < 6) {
    uint32_t idx = (uint32_t)(cls * g.n + s);
    philox_pair(key, idx, 1, 0, 0, u0, u1);
    philox_pair(key, idx, 1, 0, 1, u2, u3);
    double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
    double q0 = r0 * cos(TPI * u1), q1 = r0 * sin(TPI * u1);
    double q2 = r1 * cos(TPI * u3), q3 = r1 * sin(TPI * u3);
    double nrm = 1.0 / sqrt(q0 * q0 + q1 * q1 + q2 * q2 + q3 * q3);
    st_link(U, g, cls >> 1, cls & 1, s, su2{c_mk(q0 * nrm, q1 * nrm), c_mk(q2 * nrm, q3 * nrm)});
  } else {
    int p = cls - 6;
    uint32_t idx = (uint32_t)(p * g.n + s);
    philox_pair(key, idx, 2, 0, 0, u0, u1);
    philox_pair(key, idx, 2, 0, 1, u2, u3);
    double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
    st_higgs(A, g, p, s,
             vec3{{sigma * r0 * cos(TPI * u1), sigma * r0 * sin(TPI * u1),
                   sigma * r1 * cos(TPI * u3)}});
  }
}

void launch_hotstart(lceq_lattice* L, uint64_t seed, double sigma) {
  hipLaunchKernelGGL(k_hotstart, dim3(blocks_for(L->g.n), 8), dim3(TB), 0, L->stream, L->g, L->U,
                     L->A, Philox{(uint32_t)seed, (uint32_t)(seed >> 32)}, sigma);
}

// f/confget.f:295-307 (mode 0) and f/confget.f:528-575 on one rank (mode 1)
__global__ __launch_bounds__(256) void k_startconf(Geom g, double* __restrict__ U,
                                                   double* __restrict__ A, int mode) {
  int s = blockIdx.x * TB + threadIdx.x;
  int cls = blockIdx.y;
  if (s >= g.n) return;
  if (cls < 6) {
    int mu = cls >> 1;
    su2 u;
    if (mode == 0) {
      u = su2{c_mk(1.0, 0.0), c_mk(0.0, 0.0)};
    } else {
      double z = (double)(s / g.h2) * (mu + 1);
      u = su2{c_mk(0.0, z), c_mk(0.0, z)};
    }
    st_link(U, g, mu, cls & 1, s, u);
  } else {
    double c = 1.0 / sqrt(3.0);
    st_higgs(A, g, cls - 6, s, mode == 0 ? vec3{{0.0, 0.0, 1.0}} : vec3{{c, c, c}});
  }
}

void launch_unitconf(lceq_lattice* L) {
  hipLaunchKernelGGL(k_startconf, dim3(blocks_for(L->g.n), 8), dim3(TB), 0, L->stream, L->g, L->U,
                     L->A, 0);
}
void launch_testconf(lceq_lattice* L) {
  hipLaunchKernelGGL(k_startconf, dim3(blocks_for(L->g.n), 8), dim3(TB), 0, L->stream, L->g, L->U,
                     L->A, 1);
}

}  // namespace lceq
