// lceq_sites.hpp -- per-site update code shared by the single-lattice
// kernels (lceq_kernels.hip) and the batched-replica kernels
// (lceq_batch.hip): the staple, the Kennedy-Pendleton heatbath, link
// overrelaxation with the hopping Metropolis, the hopping force and the three
// Higgs updates, each for one site of the parity being updated.
#pragma once
#include <math.h>

#include "lceq_device.hpp"

namespace lceq {

static constexpr double TPI = 2.0 * M_PI;  // 2*acos(-1.0), f/heat_kpi.f:29

// Per-lattice parameters of one (eo, mu) link launch
struct LinkSite {
  Geom g;
  double* U;
  const double* A;
  int eo, mu;
  double beta, beta_a;
  const double* fed;  // fed uniforms of this (eo, mu) block, reference order
  Philox key;
  uint64_t pass;
  int measflag, renorm;
};

// The 12 links of the staple V(x) of f/staples.f:65-216 for x in parity eo,
// direction mu, two nu per site: t = 0, 1 for nu = the two directions != mu
// in increasing order (f/staples.f:48-52).
struct StapleLinks {
  su2 mu_xpnu[2];    // U_mu(x+nu)      parity oe  (link 3)
  su2 nu_x[2];       // U_nu(x)         parity eo  (link 2)
  su2 nu_xpmu[2];    // U_nu(x+mu)      parity oe  (link 4)
  su2 mu_xmnu[2];    // U_mu(x-nu)      parity oe  (link 6)
  su2 nu_xmnu[2];    // U_nu(x-nu)      parity oe  (link 5)
  su2 nu_xpmumnu[2]; // U_nu(x+mu-nu)   parity eo  (link 7)
};

__device__ __forceinline__ void load_staple_links(const double* __restrict__ U, const Geom& g,
                                                  int eo, int mu, unsigned s, const Steps& st,
                                                  unsigned sxm, StapleLinks& k) {
  int oe = eo ^ 1;
  int pmu = step_of(st.p, mu);
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int nu = (t == 0) ? (mu == 0 ? 1 : 0) : (mu == 2 ? 1 : 2);
    int pnu = step_of(st.p, nu), mnu = step_of(st.m, nu);
    unsigned sxn = (unsigned)((int)s + pnu);
    unsigned sxmn = (unsigned)((int)s + mnu);
    unsigned sxmmn = (unsigned)((int)s + pmu + mnu);
    k.mu_xpnu[t] = ld_link(U, g, mu, oe, sxn);
    k.nu_x[t] = ld_link(U, g, nu, eo, s);
    k.nu_xpmu[t] = ld_link(U, g, nu, oe, sxm);
    k.mu_xmnu[t] = ld_link(U, g, mu, oe, sxmn);
    k.nu_xmnu[t] = ld_link(U, g, nu, oe, sxmn);
    k.nu_xpmumnu[t] = ld_link(U, g, nu, eo, sxmmn);
  }
}

// V = sum_t [ 4 3^+ 2^+ + 7^+ 6^+ 5 ] in the accumulation order of f/staples.f:210-216,
// each outer product accumulated into V by its own multiply-add chain
__device__ __forceinline__ su2 staple_sum(const StapleLinks& k) {
  su2 V{c_mk(0.0, 0.0), c_mk(0.0, 0.0)};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    su2 up1 = m_dd(k.mu_xpnu[t], k.nu_x[t]);
    su2 bt1 = m_dn(k.mu_xmnu[t], k.nu_xmnu[t]);
    V = q_prod_acc<false, false>(V, k.nu_xpmu[t], up1);
    V = q_prod_acc<true, false>(V, k.nu_xpmumnu[t], bt1);
  }
  return V;
}

__device__ __forceinline__ su2 staple(const double* __restrict__ U, const Geom& g, int eo, int mu,
                                      unsigned s, Coord c) {
  Steps st = site_steps(g, c);
  StapleLinks k;
  load_staple_links(U, g, eo, mu, s, st, (unsigned)((int)s + step_of(st.p, mu)), k);
  return staple_sum(k);
}

// The hopping action beta_a Re[(X(a) U)(X(b) U^+)]_11 of f/agauge.f:126-159,
// which is -beta_a a.R(U) b (R as in adj_rot):
//   a.R b = (c0^2 - |c|^2)(a.b) + 2 (c.a)(c.b) - 2 c0 c.(b x a),
// equal to the products in exact arithmetic.  a.b and b x a do not depend on
// U and serve the old and the new link.
struct HopShared {
  double ab;
  double bxa[3];
};
__device__ __forceinline__ HopShared hop_shared(const vec3& a, const vec3& b) {
  return HopShared{(a.x[0] * b.x[0] + a.x[1] * b.x[1]) + a.x[2] * b.x[2],
                   {b.x[1] * a.x[2] - b.x[2] * a.x[1], b.x[2] * a.x[0] - b.x[0] * a.x[2],
                    b.x[0] * a.x[1] - b.x[1] * a.x[0]}};
}
__device__ __forceinline__ double hop_action(double beta_a, const vec3& a, const vec3& b,
                                             const HopShared& h, su2 u) {
  double c0 = u.a.re, c1 = u.b.im, c2 = u.b.re, c3 = u.a.im;
  double s = c0 * c0 - ((c1 * c1 + c2 * c2) + c3 * c3);
  double ca = (c1 * a.x[0] + c2 * a.x[1]) + c3 * a.x[2];
  double cb = (c1 * b.x[0] + c2 * b.x[1]) + c3 * b.x[2];
  double tr = (c1 * h.bxa[0] + c2 * h.bxa[1]) + c3 * h.bxa[2];
  return -beta_a * ((s * h.ab + 2.0 * (ca * cb)) - 2.0 * (c0 * tr));
}

// U / |U| (f/renorm.f:24-34) with the reciprocal square root (within an ulp
// of 1/sqrt, one operation instead of a square root and a division)
__device__ __forceinline__ su2 normalized(su2 u) {
  double t = (u.a.re * u.a.re + u.a.im * u.a.im) + (u.b.re * u.b.re + u.b.im * u.b.im);
  t = inv_sqrt(t);
  return su2{c_scale(t, u.a), c_scale(t, u.b)};
}

// Operands of one link update: the staple links, the link, the two Higgs
// triples of the hopping term and the uniforms (with the uniform-only part of
// the Kennedy-Pendleton step).
struct LinkOperands {
  StapleLinks k;
  su2 u;
  vec3 ax, axp;
  double r4, racc;                          // KP acceptance and Metropolis uniforms
  double r1, r2, r3, r5, r6;                // KP uniforms (until link_prepare)
  double kp_q, kp_a2, kp_st, kp_sin, kp_cos;  // f/heat_kpi.f:42-98, uniform-only part
};

// Load phase: the 30 operand loads (and the fed uniforms) of site s.
// EARLY: see link_update_site.
template <bool KP, bool FED, bool EARLY>
__device__ __forceinline__ void link_load(const LinkSite& p, unsigned s, LinkOperands& o) {
  const Geom& g = p.g;
  int eo = p.eo, oe = eo ^ 1, mu = p.mu;
  // The link and the Higgs pair first: the old hopping action needs only
  // them, so it is computed while the staple links are still arriving.
  constexpr bool early = !KP && EARLY;
  if (early) {
    o.u = ld_link(p.U, g, mu, eo, s);
    o.ax = ld_higgs(p.A, g, eo, s);
    __builtin_amdgcn_sched_barrier(0);
  }
  Coord c = site_coord(g, eo, (int)s);
  Steps st = site_steps(g, c);
  unsigned sxm = (unsigned)((int)s + step_of(st.p, mu));
  if (!early) {
    o.u = ld_link(p.U, g, mu, eo, s);
    o.ax = ld_higgs(p.A, g, eo, s);
  }
  o.axp = ld_higgs(p.A, g, oe, sxm);
  load_staple_links(p.U, g, eo, mu, s, st, sxm, o.k);
  if (FED) {
    const double* f = p.fed;
    unsigned n = (unsigned)g.n;
    if (KP) {
      o.r1 = f[3 * s];
      o.r2 = f[3 * s + 1];
      o.r3 = f[3 * s + 2];
      o.r4 = f[3 * n + s];
      o.r5 = f[4 * n + s];
      o.r6 = f[5 * n + s];
      o.racc = f[6 * n + s];
    } else {
      o.racc = f[s];
    }
  }
}

// Uniform phase, independent of the loads: the Philox draws and the part of
// the Kennedy-Pendleton step that depends only on them (operations and order
// of f/heat_kpi.f:42-98), so that it runs under the load latency.
template <bool KP, bool FED>
__device__ __forceinline__ void link_prepare(const LinkSite& p, unsigned s, LinkOperands& o) {
  if (!FED) {
    double dummy;
    if (KP) {
      philox_pair(p.key, s, 0, p.pass, 0, o.r1, o.r2);
      philox_pair(p.key, s, 0, p.pass, 1, o.r3, o.r4);
      philox_pair(p.key, s, 0, p.pass, 2, o.r5, o.r6);
      philox_pair(p.key, s, 0, p.pass, 3, o.racc, dummy);
    } else {
      philox_pair(p.key, s, 0, p.pass, 0, o.racc, dummy);
    }
  }
  if (KP) {
    double r1 = o.r1, r2 = o.r2;
    if (r1 == 0.0) r1 = 0.0000033;
    if (r2 == 0.0) r2 = 0.0000066;
    // cos(2 pi r) as cospi(2 r) in the product library (cos_2pi): no
    // argument reduction by a rounded 2 pi (the reference's cos(tpi*r3)
    // differs by an ulp or so); 0.5 us less per KP launch at 64^3
    double cc = cos_2pi(o.r3);
    o.kp_q = log(r1) + log(r2) * (cc * cc);
    o.kp_a2 = 2.0 * (o.r5 - 0.5);
    double tt = 1.0 - o.kp_a2 * o.kp_a2;
    o.kp_st = tt >= 0.0 ? sqrt(tt) : -1.0;
    sincos_2pi(o.r6, &o.kp_sin, &o.kp_cos);  // sin, cos(2 pi r6)
  }
}

// Update phase of one (eo, mu) class of f/agauge.f:53-193 (KP=true), or the
// link overrelaxation U -> Vhat^+ U^+ Vhat^+ with the same hopping
// Metropolis (KP=false).  Partials (added to red): 0 plaquette, 1 KP
// accepts, 2 Metropolis accepts.
template <bool KP>
__device__ __forceinline__ void link_finish(const LinkSite& p, unsigned s, const LinkOperands& o,
                                            double (&red)[3]) {
  su2 u = o.u;
  // the old hopping action (f/agauge.f:124-126, 140-147) from the first operands to
  // arrive, ahead of the staple sum; 1.7% off the overrelaxation launch
  HopShared hs = hop_shared(o.ax, o.axp);
  double sa_o = hop_action(p.beta_a, o.ax, o.axp, hs, u);
  su2 V = staple_sum(o.k);
  // f/agauge.f:71-84
  double t = (V.a.re * V.a.re + V.a.im * V.a.im) + (V.b.re * V.b.re + V.b.im * V.b.im);
  t = inv_sqrt(t);
  su2 vh{c_scale(t, V.a), c_scale(t, V.b)};
  su2 h;
  if (KP) {
    double b = t / p.beta;
    double a0old = c_mul(u.a, vh.a).re - c_mul(u.b, c_conj(vh.b)).re;
    // Kennedy-Pendleton, f/heat_kpi.f:42-98
    double a0 = -o.kp_q * b;
    bool lacc = o.r4 * o.r4 <= (1.0 - 0.5 * a0);
    red[1] += lacc ? 1.0 : 0.0;
    a0 = lacc ? 1.0 - a0 : a0old;
    double a1 = 1.0 - a0 * a0;
    a1 = a1 >= 0.0 ? sqrt(a1) : 0.0;
    su2 w;
    w.a = c_mk(a0, a1 * o.kp_a2);
    double a0b = o.kp_st >= 0.0 ? a1 * o.kp_st : 0.0;
    w.b = c_mk(a0b * o.kp_sin, a0b * o.kp_cos);
    h = m_nd(w, vh);  // f/agauge.f:98
  } else {
    // U -> Vhat^+ U^+ Vhat^+.  For a unit Vhat this is the reflection
    // 2 <U,W> W - U of U about W = Vhat^+ in the 4-vectors (Re a, Im a, Re b,
    // Im b), with <U,W> = Re(U Vhat)_11: 9 operations instead of two products
    double d2 = 2.0 * ((u.a.re * vh.a.re - u.a.im * vh.a.im) -
                       (u.b.re * vh.b.re + u.b.im * vh.b.im));
    h = su2{c_mk(d2 * vh.a.re - u.a.re, -(d2 * vh.a.im) - u.a.im),
            c_mk(-(d2 * vh.b.re) - u.b.re, -(d2 * vh.b.im) - u.b.im)};
  }
  double sa_n = hop_action(p.beta_a, o.ax, o.axp, hs, h);
  if (exp(sa_o - sa_n) >= o.racc) {  // f/agauge.f:167-174
    u = p.renorm ? normalized(h) : h;
    st_link(p.U, p.g, p.mu, p.eo, s, u);
    red[2] += 1.0;
  }
  if (p.measflag == 1) red[0] += m_nn(u, V).a.re;  // f/agauge.f:179-185
}

// One site: all operand loads are issued before any arithmetic, so that one
// memory latency covers the site; the uniform phase runs under it.
// EARLY selects the schedule for launches of many waves whose fields sit in
// the Infinity Cache (the launchers pick it per lattice):
// - overrelaxation issues the link and the Higgs triple at the site itself
//   ahead of the neighbour index arithmetic (the OR launch 2.4% faster at
//   64^3, 2.2% at 128^3; 1.4% slower at 160^3, where HBM bounds it; the
//   heatbath was slower with it);
// - a fence keeps the uniform phase ahead of the update arithmetic: without
//   it the compiler sinks part of the Philox rounds below the last operand
//   wait, into the launch's exposed tail (KP 7.13 -> 6.91 us, OR 5.13 ->
//   5.00 us at 64^3).
// Launches of fewer than 512 waves (below 40^3), where one wave's chain is
// the launch, keep the previous schedule: alternated A/B runs there were
// within their 10% run-to-run spread.  The fields are the same bits either
// way.
template <bool KP, bool FED, bool EARLY = true>
__device__ __forceinline__ void link_update_site(const LinkSite& p, unsigned s, double (&red)[3]) {
  LinkOperands o;
  link_load<KP, FED, EARLY>(p, s, o);
  __builtin_amdgcn_sched_barrier(0);
  link_prepare<KP, FED>(p, s, o);
  if constexpr (EARLY) __builtin_amdgcn_sched_barrier(0);
  link_finish<KP>(p, s, o, red);
}

// Operands of the hopping force F(x) of f/a_force.f:54-110 for x in parity eo
struct ForceLinks {
  su2 ux[3];   // U_mu(x)        parity eo
  su2 uy[3];   // U_mu(x-mu)     parity oe
  vec3 axp[3]; // a(x+mu)        parity oe
  vec3 axd[3]; // a(x-mu)        parity oe
};

// the neighbour operands of the force (the caller loads the links U_mu(x)
// at the site itself)
__device__ __forceinline__ void load_force_neighbours(const double* __restrict__ U,
                                                      const double* __restrict__ A,
                                                      const Geom& g, int eo, unsigned s, Coord c,
                                                      ForceLinks& k) {
  int oe = eo ^ 1;
  Steps st = site_steps(g, c);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    unsigned sxp = (unsigned)((int)s + st.p[mu]), sxd = (unsigned)((int)s + st.m[mu]);
    k.uy[mu] = ld_link(U, g, mu, oe, sxd);
    k.axp[mu] = ld_higgs(A, g, oe, sxp);
    k.axd[mu] = ld_higgs(A, g, oe, sxd);
  }
}

__device__ __forceinline__ void load_force_links(const double* __restrict__ U,
                                                 const double* __restrict__ A, const Geom& g,
                                                 int eo, unsigned s, Coord c, ForceLinks& k) {
  int oe = eo ^ 1;
  Steps st = site_steps(g, c);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    unsigned sxp = (unsigned)((int)s + st.p[mu]), sxd = (unsigned)((int)s + st.m[mu]);
    k.ux[mu] = ld_link(U, g, mu, eo, s);
    k.uy[mu] = ld_link(U, g, mu, oe, sxd);
    k.axp[mu] = ld_higgs(A, g, oe, sxp);
    k.axd[mu] = ld_higgs(A, g, oe, sxd);
  }
}

// F = sum_mu extract[U(x) X(a(x+mu)) U(x)^+ + U(x-mu)^+ X(a(x-mu)) U(x-mu)],
// each term as the rotation adj_rot
__device__ __forceinline__ vec3 force_sum(const ForceLinks& k) {
  vec3 F{{0.0, 0.0, 0.0}};
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    vec3 t1 = adj_rot<false>(k.ux[mu], k.axp[mu]);
    vec3 t2 = adj_rot<true>(k.uy[mu], k.axd[mu]);
#pragma unroll
    for (int i = 0; i < 3; ++i) F.x[i] = F.x[i] + (t1.x[i] + t2.x[i]);
  }
  return F;
}

__device__ __forceinline__ vec3 hop_force(const double* __restrict__ U,
                                          const double* __restrict__ A, const Geom& g, int eo,
                                          unsigned s, Coord c) {
  ForceLinks k;
  load_force_links(U, A, g, eo, s, c, k);
  return force_sum(k);
}

// Per-lattice parameters of one Higgs parity launch
struct HiggsSite {
  Geom g;
  const double* U;
  double* A;
  int eo, imode;
  double beta_a, beta_2, beta_4;
  const double* delta;  // device scalar (ahiggs1)
  const double* fed;
  Philox key;
  uint64_t pass;
  int measflag;
};

// KIND 1: f/ahiggs1.f Metropolis; 2: f/ahiggs2.f heatbath; 3: f/ahiggs3.f
// overrelaxation.  Partials: 0 accepts, 1 rho, 2 rho^2, 3 a.F (post update).
// EARLY (ahiggs1 only): the operands at the site itself go out ahead of the
// neighbour index arithmetic, as in link_update_site (1.3% off the ahiggs1
// launch at 64^3; ahiggs2 and ahiggs3 were slower with it).
template <int KIND, bool FED, bool EARLY = true>
__device__ __forceinline__ void higgs_update_site(const HiggsSite& p, unsigned s, double (&red)[4]) {
  const Geom& g = p.g;
  int eo = p.eo;
  unsigned n = (unsigned)g.n;
  // ---- load phase: 6 links, 7 Higgs triples and the fed uniforms
  ForceLinks k;
  vec3 a, an;
  if (KIND == 1 && EARLY) {
    a = ld_higgs(p.A, g, eo, s);
#pragma unroll
    for (int mu = 0; mu < 3; ++mu) k.ux[mu] = ld_link(p.U, g, mu, eo, s);
    __builtin_amdgcn_sched_barrier(0);
    load_force_neighbours(p.U, p.A, g, eo, s, site_coord(g, eo, (int)s), k);
  } else {
    load_force_links(p.U, p.A, g, eo, s, site_coord(g, eo, (int)s), k);
    a = ld_higgs(p.A, g, eo, s);
  }
  double fu[7];
  // ahiggs1's step size lives on the device (it self-tunes there); its load
  // goes out with the site's operand loads, not after them
  double delta = KIND == 1 ? *p.delta : 0.0;
  if (FED) {
    if (KIND == 1) {
      fu[0] = p.fed[3 * s];
      fu[1] = p.fed[3 * s + 1];
      fu[2] = p.fed[3 * s + 2];
      fu[3] = p.fed[3 * n + s];
    } else if (KIND == 2) {
#pragma unroll
      for (int q = 0; q < 7; ++q) fu[q] = p.fed[(size_t)q * n + s];
    } else {
      fu[0] = p.fed[s];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  vec3 F = force_sum(k);
  double ro = 0.0, rn = 0.0, boltz, racc;
  if (KIND == 1) {  // f/ahiggs1.f:51-92
    double r[3];
    if (FED) {
      r[0] = fu[0];
      r[1] = fu[1];
      r[2] = fu[2];
      racc = fu[3];
    } else {
      philox_pair(p.key, (uint32_t)s, 0, p.pass, 0, r[0], r[1]);
      double dummy;
      philox_pair(p.key, (uint32_t)s, 0, p.pass, 1, r[2], racc);
      (void)dummy;
    }
    double sfo = 0.0, sfn = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      an.x[i] = a.x[i] + (r[i] - 0.5) * delta;
      sfo = sfo + a.x[i] * F.x[i];
      sfn = sfn + an.x[i] * F.x[i];
      ro = ro + a.x[i] * a.x[i];
      rn = rn + an.x[i] * an.x[i];
    }
    boltz = exp(-(p.beta_a * (sfo - sfn)) + p.beta_2 * (ro - rn) +
                p.beta_4 * (ro * ro - rn * rn));
  } else {
    double alph, width, cx;
    if (p.imode == 1) {  // f/ahiggs2.f:61-77
      double conf2 = 0.0;
#pragma unroll
      for (int i = 0; i < 3; ++i) conf2 = conf2 + p.beta_a * (F.x[i] * F.x[i]);
      double d = p.beta_2 + p.beta_4 * conf2;
      alph = (p.beta_4 * conf2) / (2.0 * (d * d));
      width = p.beta_2 + alph;
      cx = 1.0 / 2.0 / width;
    } else {  // f/ahiggs2.f:52-56
      alph = 0.0;
      width = p.beta_2;
      cx = 1.0 / p.beta_2 / 2.0;
    }
    if (KIND == 2) {  // gauss x3 (f/gauss.f:18-29), f/ahiggs2.f:83-91
      double w[6];
      if (FED) {
#pragma unroll
        for (int q = 0; q < 6; ++q) w[q] = fu[q];
        racc = fu[6];
      } else {
        double dummy;
        philox_pair(p.key, (uint32_t)s, 0, p.pass, 0, w[0], w[1]);
        philox_pair(p.key, (uint32_t)s, 0, p.pass, 1, w[2], w[3]);
        philox_pair(p.key, (uint32_t)s, 0, p.pass, 2, w[4], w[5]);
        philox_pair(p.key, (uint32_t)s, 0, p.pass, 3, racc, dummy);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        double w1 = -1.0 * (1.0 / width) * log(w[2 * i]);
        double gi = sqrt(w1) * cos_2pi(w[2 * i + 1]);  // cos(2 pi w), as in the KP step
        an.x[i] = gi + p.beta_a * F.x[i] * cx;
      }
    } else {  // f/ahiggs3.f:83-87
      if (FED) {
        racc = fu[0];
      } else {
        double dummy;
        philox_pair(p.key, (uint32_t)s, 0, p.pass, 0, racc, dummy);
      }
#pragma unroll
      for (int i = 0; i < 3; ++i) an.x[i] = -a.x[i] + 2.0 * p.beta_a * F.x[i] * cx;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ro = ro + a.x[i] * a.x[i];
      rn = rn + an.x[i] * an.x[i];
    }
    boltz = exp(-(alph * (ro - rn)) + p.beta_4 * (ro * ro - rn * rn));
  }
  if (boltz >= racc) {
    a = an;
    ro = rn;
    st_higgs(p.A, g, eo, s, a);
    red[0] = 1.0;
  }
  if (p.measflag == 1) {
    red[1] = ro;
    red[2] = ro * ro;
    red[3] = (a.x[0] * F.x[0] + a.x[1] * F.x[1]) + a.x[2] * F.x[2];
  }
}

}  // namespace lceq
