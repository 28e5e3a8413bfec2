/*
 * lceq_oracle.c -- CPU restatement (FP64) of the lceqcd lattice update and
 * measurement path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the timed
 * CPU baseline.  The product library never links or calls this file.
 *
 * Every routine restates one reference subroutine loop by loop, in the same
 * order of floating-point operations where the reference fixes one, and cites
 * the file:line it follows.  Uniforms come either from a fed buffer consumed
 * sequentially in the reference's own call order (each random_number call
 * site of the reference becomes one oc_draw() here, in the same loop nest), or
 * from the counter-based Philox4x32-10 streams that the product uses in
 * production (documented at philox_uniform below).
 *
 * Deliberate deviation: the gauge-fixing angle acos(Re Omega_11)
 * (f/or_gauge_fixing1.f:174) clamps its argument to [-1,1]; the reference
 * returns NaN when rounding pushes |Re Omega_11| one ulp above 1.
 */
#include "lceq_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  double re, im;
} cpx;

static inline cpx c_mk(double r, double i) {
  cpx z = {r, i};
  return z;
}
static inline cpx c_add(cpx a, cpx b) { return c_mk(a.re + b.re, a.im + b.im); }
static inline cpx c_sub(cpx a, cpx b) { return c_mk(a.re - b.re, a.im - b.im); }
static inline cpx c_neg(cpx a) { return c_mk(-a.re, -a.im); }
static inline cpx c_conj(cpx a) { return c_mk(a.re, -a.im); }
static inline cpx c_mul(cpx a, cpx b) {
  return c_mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cpx c_scale(double t, cpx a) { return c_mk(t * a.re, t * a.im); }

/* a link or algebra element in (alpha, beta) form, f/INCLUDE/gaugefield.inc:89-101 */
typedef struct {
  cpx a, b;
} su2;

static inline su2 ld_su2(const double *p) {
  su2 u;
  u.a = c_mk(p[0], p[1]);
  u.b = c_mk(p[2], p[3]);
  return u;
}
static inline void st_su2(double *p, su2 u) {
  p[0] = u.a.re;
  p[1] = u.a.im;
  p[2] = u.b.re;
  p[3] = u.b.im;
}

/* f/matmul.f:22-23  U1*U2 */
static inline su2 m_nn(su2 p, su2 q) {
  su2 h;
  h.a = c_sub(c_mul(p.a, q.a), c_mul(p.b, c_conj(q.b)));
  h.b = c_add(c_mul(p.a, q.b), c_mul(p.b, c_conj(q.a)));
  return h;
}
/* f/matmul.f:52-56  U1^+ * U2 */
static inline su2 m_dn(su2 p, su2 q) {
  su2 h;
  h.a = c_add(c_mul(c_conj(p.a), q.a), c_mul(p.b, c_conj(q.b)));
  h.b = c_sub(c_mul(c_conj(p.a), q.b), c_mul(p.b, c_conj(q.a)));
  return h;
}
/* f/matmul.f:85-89  U1 * U2^+ */
static inline su2 m_nd(su2 p, su2 q) {
  su2 h;
  h.a = c_add(c_mul(p.a, c_conj(q.a)), c_mul(p.b, c_conj(q.b)));
  h.b = c_add(c_neg(c_mul(p.a, q.b)), c_mul(p.b, q.a));
  return h;
}
/* f/matmul.f:118-122  U1^+ * U2^+ */
static inline su2 m_dd(su2 p, su2 q) {
  su2 h;
  h.a = c_sub(c_mul(c_conj(p.a), c_conj(q.a)), c_mul(p.b, c_conj(q.b)));
  h.b = c_sub(c_neg(c_mul(c_conj(p.a), q.b)), c_mul(p.b, q.a));
  return h;
}
/* f/matmul.f:150-154  X(a) * U^+ with X(a) = (i a3, a2 + i a1) */
static inline su2 m_ad(const double *a, su2 u) {
  su2 h;
  cpx x1 = c_mk(0.0, a[2]), x2 = c_mk(a[1], a[0]);
  h.a = c_add(c_mul(x1, c_conj(u.a)), c_mul(x2, c_conj(u.b)));
  h.b = c_add(c_neg(c_mul(x1, u.b)), c_mul(x2, u.a));
  return h;
}
/* f/matmul.f:183-187  X(a) * U */
static inline su2 m_an(const double *a, su2 u) {
  su2 h;
  cpx x1 = c_mk(0.0, a[2]), x2 = c_mk(a[1], a[0]);
  h.a = c_sub(c_mul(x1, u.a), c_mul(x2, c_conj(u.b)));
  h.b = c_add(c_mul(x1, u.b), c_mul(x2, c_conj(u.a)));
  return h;
}

struct oc_lattice {
  int lx, ly, lz;
  int h1, h2, n, v; /* ownhalfvol(1..3), ownvol(3) */
  int32_t *up, *dn; /* siteup/sitedn, 0-based, [(p*3+mu)*n + s] */
  double *u, *a, *uf, *af, *stap, *force;
  double beta, beta_a, beta_2, beta_4;
  double lambda, alpha, limit_or;
  int max_iter, imode;
  int renorm_on_write; /* production sweeps: normalise accepted links */
  int rng_mode;
  const double *fed;
  int64_t fed_n, fed_pos;
  uint64_t seed, pass;
  double plaq0, a0kin, a02, a04, a0f[3];
  int err;
};

/* ------------------------------------------------------------------ RNG */

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: ten rounds of the 4x32 S-box
 *   (c0,c1,c2,c3) -> (hi(M1 c2)^c1^k0, lo(M1 c2), hi(M0 c0)^c3^k1, lo(M0 c0))
 * with the key bumped by the Weyl constants between rounds. */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += W0;
    k1 += W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Production uniform stream.  Draw k of lattice index idx in pass `pass`:
 * counter = (idx, (stream<<8) | k/2, pass_lo, pass_hi), key = seed.  Words
 * (2(k%2), 2(k%2)+1) of the block form a 52-bit integer x and the uniform is
 * (2x+1) 2^-53, which lies strictly inside (0,1). */
static double philox_uniform(uint64_t seed, uint32_t idx, uint32_t stream,
                             uint64_t pass, int k) {
  uint32_t ctr[4] = {idx, (stream << 8) | (uint32_t)(k >> 1), (uint32_t)pass,
                     (uint32_t)(pass >> 32)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t w[4];
  oc_philox4x32_10(ctr, key, w);
  uint32_t hi = w[2 * (k & 1)], lo = w[2 * (k & 1) + 1];
  uint64_t x = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return (double)(2 * x + 1) * 0x1p-53;
}

/* one call of the reference's random_number (fed mode only) */
static double oc_draw(oc_lattice *L) {
  if (L->fed_pos >= L->fed_n) {
    L->err = 1;
    return 0.5;
  }
  return L->fed[L->fed_pos++];
}

/* uniform for site s, draw k of the current pass (philox mode) */
static inline double oc_pu(oc_lattice *L, uint64_t pass, int s, int k) {
  return philox_uniform(L->seed, (uint32_t)s, 0, pass, k);
}

/* ------------------------------------------------------------- geometry */

/* restatement of f/adresses.f:91-337, single process in every direction */
void oc_site_tables(const oc_lattice *L, int32_t *siteup, int32_t *sitedn) {
  int h1 = L->h1, h2 = L->h2, n = L->n;
  int i0 = 0, i1 = 0;
  for (int z = 1; z <= L->lz; ++z)
    for (int y = 1; y <= L->ly; ++y)
      for (int x = 1; x <= L->lx; ++x) {
        int eo, i;
        if ((x + y + z) % 2 == 1) {
          eo = 0;
          i = ++i0;
        } else {
          eo = 1;
          i = ++i1;
        }
        int32_t *up0 = siteup + (eo * 3 + 0) * n, *dn0 = sitedn + (eo * 3 + 0) * n;
        int32_t *up1 = siteup + (eo * 3 + 1) * n, *dn1 = sitedn + (eo * 3 + 1) * n;
        int32_t *up2 = siteup + (eo * 3 + 2) * n, *dn2 = sitedn + (eo * 3 + 2) * n;
        /* x direction, f/adresses.f:103-116 (eo=0) and :221-234 (eo=1) */
        int xshift = (eo == 0) ? ((y + z) % 2 == 1) : ((y + z) % 2 == 0);
        if (xshift) {
          up0[i - 1] = (i % h1) + 1 + ((i - 1) / h1) * h1;
          dn0[i - 1] = i;
        } else {
          up0[i - 1] = i;
          dn0[i - 1] = ((i + h1 - 2) % h1) + 1 + ((i - 1) / h1) * h1;
        }
        /* y direction, f/adresses.f:148-157 */
        up1[i - 1] = ((i - 1 + h1) % h2) + 1 + ((i - 1) / h2) * h2;
        dn1[i - 1] = ((i - 1 + h1 * (L->ly - 1)) % h2) + 1 + ((i - 1) / h2) * h2;
        /* z direction, f/adresses.f:183-193 */
        up2[i - 1] = ((i - 1 + h2) % n) + 1 + ((i - 1) / n) * n;
        dn2[i - 1] = ((i - 1 + h2 * (L->lz - 1)) % n) + 1 + ((i - 1) / n) * n;
      }
  for (int k = 0; k < 6 * n; ++k) {
    siteup[k] -= 1;
    sitedn[k] -= 1;
  }
}

/* closed form: is = floor(x0/2) + (lx/2) y0 + (lx ly/2) z0, parity of
 * x0+y0+z0; the same arithmetic the device kernels use */
void oc_site_tables_closed_form(const oc_lattice *L, int32_t *siteup,
                                int32_t *sitedn) {
  int lx = L->lx, ly = L->ly, lz = L->lz, h1 = L->h1, n = L->n;
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < n; ++s) {
      int z0 = s / (h1 * ly), r = s % (h1 * ly), y0 = r / h1, xh = r % h1;
      int x0 = 2 * xh + ((p + y0 + z0) & 1);
      for (int mu = 0; mu < 3; ++mu) {
        int c[3] = {x0, y0, z0}, d[3] = {x0, y0, z0}, ext[3] = {lx, ly, lz};
        c[mu] = (c[mu] + 1) % ext[mu];
        d[mu] = (d[mu] - 1 + ext[mu]) % ext[mu];
        siteup[(p * 3 + mu) * n + s] = c[0] / 2 + h1 * c[1] + h1 * ly * c[2];
        sitedn[(p * 3 + mu) * n + s] = d[0] / 2 + h1 * d[1] + h1 * ly * d[2];
      }
    }
}

/* ------------------------------------------------------------ lifecycle */

oc_lattice *oc_create(int lx, int ly, int lz) {
  if (lx < 2 || ly < 2 || lz < 2 || (lx | ly | lz) & 1) return NULL;
  oc_lattice *L = (oc_lattice *)calloc(1, sizeof(oc_lattice));
  L->lx = lx;
  L->ly = ly;
  L->lz = lz;
  L->h1 = lx / 2;
  L->h2 = lx * ly / 2;
  L->v = lx * ly * lz;
  L->n = L->v / 2;
  size_t n = (size_t)L->n;
  L->up = (int32_t *)malloc(6 * n * sizeof(int32_t));
  L->dn = (int32_t *)malloc(6 * n * sizeof(int32_t));
  oc_site_tables(L, L->up, L->dn);
  L->u = (double *)calloc(24 * n, sizeof(double));
  L->uf = (double *)calloc(24 * n, sizeof(double));
  L->stap = (double *)calloc(24 * n, sizeof(double));
  L->a = (double *)calloc(6 * n, sizeof(double));
  L->af = (double *)calloc(6 * n, sizeof(double));
  L->force = (double *)calloc(3 * n, sizeof(double));
  L->beta = 9.0;
  L->lambda = 1.0;
  L->alpha = 1.0;
  L->limit_or = 1e-9;
  L->max_iter = 1000000;
  L->rng_mode = OC_RNG_PHILOX;
  L->seed = 12345;
  return L;
}

void oc_destroy(oc_lattice *L) {
  if (!L) return;
  free(L->up);
  free(L->dn);
  free(L->u);
  free(L->uf);
  free(L->stap);
  free(L->a);
  free(L->af);
  free(L->force);
  free(L);
}

int oc_error(const oc_lattice *L) { return L->err; }
void oc_set_couplings(oc_lattice *L, double beta, double beta_a, double beta_2,
                      double beta_4) {
  L->beta = beta;
  L->beta_a = beta_a;
  L->beta_2 = beta_2;
  L->beta_4 = beta_4;
}
void oc_set_gauge_fixing(oc_lattice *L, double lambda, double alpha,
                         double limit_or, int max_iter) {
  L->lambda = lambda;
  L->alpha = alpha;
  L->limit_or = limit_or;
  L->max_iter = max_iter;
}
void oc_set_imode(oc_lattice *L, int imode) { L->imode = imode; }
void oc_set_renorm_on_write(oc_lattice *L, int flag) { L->renorm_on_write = flag; }
void oc_rng_feed(oc_lattice *L, const double *buf, int64_t n) {
  L->rng_mode = OC_RNG_FED;
  L->fed = buf;
  L->fed_n = n;
  L->fed_pos = 0;
}
int64_t oc_rng_consumed(const oc_lattice *L) { return L->fed_pos; }
void oc_rng_philox(oc_lattice *L, uint64_t seed, uint64_t pass_counter) {
  L->rng_mode = OC_RNG_PHILOX;
  L->seed = seed;
  L->pass = pass_counter;
}
uint64_t oc_pass_counter(const oc_lattice *L) { return L->pass; }

double *oc_links(oc_lattice *L) { return L->u; }
double *oc_higgs(oc_lattice *L) { return L->a; }
double *oc_links_fixed(oc_lattice *L) { return L->uf; }
double *oc_higgs_fixed(oc_lattice *L) { return L->af; }
double *oc_staple(oc_lattice *L) { return L->stap; }
double *oc_force(oc_lattice *L) { return L->force; }

#define UPTR(L, base, mu, p, s) ((base) + ((((size_t)(mu)*2 + (p)) * (L)->n + (s)) * 4))
#define APTR(L, base, p, s) ((base) + (((size_t)(p) * (L)->n + (s)) * 3))
#define SUP(L, p, mu, s) ((L)->up[((p)*3 + (mu)) * (L)->n + (s)])
#define SDN(L, p, mu, s) ((L)->dn[((p)*3 + (mu)) * (L)->n + (s)])

/* f/confget.f:295-307 */
void oc_unitconf(oc_lattice *L) {
  for (int mu = 0; mu < 3; ++mu)
    for (int p = 0; p < 2; ++p)
      for (int s = 0; s < L->n; ++s) {
        double *q = UPTR(L, L->u, mu, p, s);
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
      }
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < L->n; ++s) {
      double *q = APTR(L, L->a, p, s);
      q[0] = q[1] = 0.0;
      q[2] = 1.0;
    }
}

/* f/confget.f:528-575 on one rank: u_mu = mu*(i z, i z), a = (1,1,1)/sqrt(3) */
void oc_testconf(oc_lattice *L) {
  for (int mu = 0; mu < 3; ++mu)
    for (int p = 0; p < 2; ++p)
      for (int s = 0; s < L->n; ++s) {
        double z = (double)(s / L->h2);
        double *q = UPTR(L, L->u, mu, p, s);
        q[0] = 0.0 * (mu + 1);
        q[1] = z * (mu + 1);
        q[2] = 0.0 * (mu + 1);
        q[3] = z * (mu + 1);
      }
  double c = 1. / sqrt(3.);
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < L->n; ++s) {
      double *q = APTR(L, L->a, p, s);
      q[0] = q[1] = q[2] = c;
    }
}

/* Builder-defined hot start: Haar links from four Box-Muller normals
 * (stream 1), Higgs components N(0, sigma^2) (stream 2). */
void oc_hotstart(oc_lattice *L, uint64_t seed, double sigma) {
  const double tpi = 2.0 * acos(-1.0);
  for (int mu = 0; mu < 3; ++mu)
    for (int p = 0; p < 2; ++p)
      for (int s = 0; s < L->n; ++s) {
        uint32_t idx = (uint32_t)((mu * 2 + p) * L->n + s);
        double u0 = philox_uniform(seed, idx, 1, 0, 0);
        double u1 = philox_uniform(seed, idx, 1, 0, 1);
        double u2 = philox_uniform(seed, idx, 1, 0, 2);
        double u3 = philox_uniform(seed, idx, 1, 0, 3);
        double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
        double g[4] = {r0 * cos(tpi * u1), r0 * sin(tpi * u1), r1 * cos(tpi * u3),
                       r1 * sin(tpi * u3)};
        double nrm = 1.0 / sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3]);
        double *q = UPTR(L, L->u, mu, p, s);
        for (int c = 0; c < 4; ++c) q[c] = g[c] * nrm;
      }
  for (int p = 0; p < 2; ++p)
    for (int s = 0; s < L->n; ++s) {
      uint32_t idx = (uint32_t)(p * L->n + s);
      double u0 = philox_uniform(seed, idx, 2, 0, 0);
      double u1 = philox_uniform(seed, idx, 2, 0, 1);
      double u2 = philox_uniform(seed, idx, 2, 0, 2);
      double u3 = philox_uniform(seed, idx, 2, 0, 3);
      double r0 = sqrt(-2.0 * log(u0)), r1 = sqrt(-2.0 * log(u2));
      double *q = APTR(L, L->a, p, s);
      q[0] = sigma * r0 * cos(tpi * u1);
      q[1] = sigma * r0 * sin(tpi * u1);
      q[2] = sigma * r1 * cos(tpi * u3);
    }
}

/* ------------------------------------------------------------- multall */

void oc_multall(int kind, const double *op1, const double *op2, double *out,
                int n, const int32_t *s1, const int32_t *s2) {
  for (int i = 0; i < n; ++i) {
    int i1 = s1 ? s1[i] : i, i2 = s2 ? s2[i] : i;
    su2 q = ld_su2(op2 + 4 * (size_t)i2), h;
    if (kind == 4 || kind == 5) {
      const double *a = op1 + 3 * (size_t)i1;
      h = (kind == 4) ? m_ad(a, q) : m_an(a, q);
    } else {
      su2 p = ld_su2(op1 + 4 * (size_t)i1);
      switch (kind) {
        case 0: h = m_nn(p, q); break;
        case 1: h = m_dn(p, q); break;
        case 2: h = m_nd(p, q); break;
        default: h = m_dd(p, q); break;
      }
    }
    st_su2(out + 4 * (size_t)i, h);
  }
}

/* -------------------------------------------------------------- staples */

/* f/staples.f:9-226.  V(x) = sum_{nu != mu} [ U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
 *                              + U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu) ]     */
void oc_staples(oc_lattice *L, int eo, int mu) {
  int oe = 1 - eo, n = L->n;
  su2 *bt1 = (su2 *)malloc(n * sizeof(su2)), *bt2 = (su2 *)malloc(n * sizeof(su2));
  su2 *up1 = (su2 *)malloc(n * sizeof(su2)), *up2 = (su2 *)malloc(n * sizeof(su2));
  int nus[2], k = 0;
  for (int nu = 0; nu < 3; ++nu)
    if (nu != mu) nus[k++] = nu; /* f/staples.f:48-52 */
  for (int s = 0; s < n; ++s) /* f/staples.f:58-61 */
    memset(UPTR(L, L->stap, mu, eo, s), 0, 4 * sizeof(double));
  for (int t = 0; t < 2; ++t) {
    int nu = nus[t];
    for (int y = 0; y < n; ++y) /* :99  6^+ 5 */
      bt1[y] = m_dn(ld_su2(UPTR(L, L->u, mu, oe, y)), ld_su2(UPTR(L, L->u, nu, oe, y)));
    for (int x = 0; x < n; ++x) /* :130  3^+ 2^+ */
      up1[x] = m_dd(ld_su2(UPTR(L, L->u, mu, oe, SUP(L, eo, nu, x))),
                    ld_su2(UPTR(L, L->u, nu, eo, x)));
    for (int y = 0; y < n; ++y) /* :162  7^+ (6^+ 5) */
      bt2[y] = m_dn(ld_su2(UPTR(L, L->u, nu, eo, SUP(L, oe, mu, y))), bt1[y]);
    for (int x = 0; x < n; ++x) /* :194  4 (3^+ 2^+) */
      up2[x] = m_nn(ld_su2(UPTR(L, L->u, nu, oe, SUP(L, eo, mu, x))), up1[x]);
    for (int x = 0; x < n; ++x) { /* :210-216 */
      double *vp = UPTR(L, L->stap, mu, eo, x);
      su2 v = ld_su2(vp), b = bt2[SDN(L, eo, nu, x)];
      v.a = c_add(c_add(v.a, up2[x].a), b.a);
      v.b = c_add(c_add(v.b, up2[x].b), b.b);
      st_su2(vp, v);
    }
  }
  free(bt1);
  free(bt2);
  free(up1);
  free(up2);
}

/* ------------------------------------------------------------- heat_kpi */

/* f/heat_kpi.f:9-101 (heat_kpi).  In philox mode draw k of site s is
 * r1,r2,r3,r4,r5,r6 for k = 0..5. */
static void heat_kpi_core(oc_lattice *L, int n, const double *b, su2 *v,
                          const double *a0_old, int64_t *nacch, uint64_t pass) {
  const double tpi = 2.0 * acos(-1.0);
  double *a0 = (double *)malloc(n * sizeof(double)), *a1 = (double *)malloc(n * sizeof(double));
  double *a2 = (double *)malloc(n * sizeof(double)), *a3 = (double *)malloc(n * sizeof(double));
  int fed = L->rng_mode == OC_RNG_FED;
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:33-40 */
    a1[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 0);
    a2[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 1);
    a3[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 2);
  }
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:42-49 */
    if (a1[s] == 0.0) a1[s] = 0.0000033;
    if (a2[s] == 0.0) a2[s] = 0.0000066;
    double c = cos(tpi * a3[s]);
    a0[s] = -(log(a1[s]) + log(a2[s]) * (c * c)) * b[s];
  }
  for (int s = 0; s < n; ++s) a1[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 3); /* f/heat_kpi.f:53-56 */
  int64_t cnt = 0;
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:58-64 */
    int lacc = a1[s] * a1[s] <= (1.0 - 0.5 * a0[s]);
    cnt += lacc;
    a0[s] = lacc ? 1.0 - a0[s] : a0_old[s];
  }
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:68-73 */
    a1[s] = 1.0 - a0[s] * a0[s];
    a1[s] = a1[s] >= 0.0 ? sqrt(a1[s]) : 0.0;
  }
  for (int s = 0; s < n; ++s) a2[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 4); /* f/heat_kpi.f:75-78 */
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:80-88 */
    a2[s] = 2.0 * (a2[s] - 0.5);
    a3[s] = a1[s] * a2[s];
    v[s].a = c_mk(a0[s], a3[s]);
    double t = 1.0 - a2[s] * a2[s];
    a0[s] = t >= 0.0 ? a1[s] * sqrt(t) : 0.0;
  }
  for (int s = 0; s < n; ++s) a3[s] = fed ? oc_draw(L) : oc_pu(L, pass, s, 5); /* f/heat_kpi.f:90-93 */
  for (int s = 0; s < n; ++s) { /* f/heat_kpi.f:95-98 */
    double ang = tpi * a3[s];
    double c1 = a0[s] * cos(ang), c2 = a0[s] * sin(ang);
    v[s].b = c_mk(c2, c1);
  }
  *nacch = cnt;
  free(a0);
  free(a1);
  free(a2);
  free(a3);
}

void oc_heat_kpi(oc_lattice *L, int n, const double *b, double *v,
                 const double *a0_old, int64_t *nacch) {
  su2 *w = (su2 *)malloc(n * sizeof(su2));
  uint64_t pass = L->pass;
  if (L->rng_mode == OC_RNG_PHILOX) L->pass += 1;
  heat_kpi_core(L, n, b, w, a0_old, nacch, pass);
  for (int s = 0; s < n; ++s) st_su2(v + 4 * (size_t)s, w[s]);
  free(w);
}

static void renorm_field(double *u, size_t nlinks);

/* ----------------------------------------------------- hop term helpers */

/* beta_A Re[(X(a_x) U)(X(a_{x+mu}) U^+)]_11, f/agauge.f:126-159 */
static inline double hop_action(double beta_a, const double *ax, const double *axp,
                                su2 u) {
  su2 t1 = m_an(ax, u), t3 = m_ad(axp, u);
  su2 t4 = m_nn(t1, t3);
  return beta_a * t4.a.re;
}

/* ------------------------------------------------------- link updates */

/* f/agauge.f:7-205 (Kennedy-Pendleton heatbath + Higgs-hopping Metropolis).
 * kp=1: agauge; kp=0: link overrelaxation U -> Vhat^+ U^+ Vhat^+ (no
 * reference counterpart) with the same Metropolis step. */
static void link_update(oc_lattice *L, int kp, int measflag, double *accept) {
  int n = L->n, fed = L->rng_mode == OC_RNG_FED;
  double *b = (double *)malloc(n * sizeof(double));
  double *a0o = (double *)malloc(n * sizeof(double));
  double *sa_o = (double *)malloc(n * sizeof(double));
  double *sa_n = (double *)malloc(n * sizeof(double));
  su2 *vh = (su2 *)malloc(n * sizeof(su2)), *w = (su2 *)malloc(n * sizeof(su2));
  su2 *h = (su2 *)malloc(n * sizeof(su2));
  double acc_kp = 0.0, acc1 = 0.0;
  L->plaq0 = 0.0;
  for (int eo = 0; eo < 2; ++eo) {
    int oe = 1 - eo;
    for (int mu = 0; mu < 3; ++mu) {
      uint64_t pass = L->pass;
      if (!fed) L->pass += 1;
      oc_staples(L, eo, mu);
      for (int s = 0; s < n; ++s) { /* f/agauge.f:71-84 */
        su2 v = ld_su2(UPTR(L, L->stap, mu, eo, s));
        su2 u = ld_su2(UPTR(L, L->u, mu, eo, s));
        double t = (v.a.re * v.a.re + v.a.im * v.a.im) + (v.b.re * v.b.re + v.b.im * v.b.im);
        t = 1.0 / sqrt(t);
        vh[s].a = c_scale(t, v.a);
        vh[s].b = c_scale(t, v.b);
        b[s] = t / L->beta;
        a0o[s] = c_mul(u.a, vh[s].a).re - c_mul(u.b, c_conj(vh[s].b)).re;
      }
      if (kp) {
        int64_t nacch;
        heat_kpi_core(L, n, b, w, a0o, &nacch, pass); /* :93 */
        acc_kp += (double)nacch;
        for (int s = 0; s < n; ++s) h[s] = m_nd(w[s], vh[s]); /* :98-104 */
      } else {
        for (int s = 0; s < n; ++s) {
          su2 u = ld_su2(UPTR(L, L->u, mu, eo, s));
          su2 r = m_nn(m_nn(vh[s], u), vh[s]); /* Vhat U Vhat */
          h[s].a = c_conj(r.a);                /* (Vhat U Vhat)^+ */
          h[s].b = c_neg(r.b);
        }
      }
      for (int s = 0; s < n; ++s) { /* :126-159 */
        const double *ax = APTR(L, L->a, eo, s), *axp = APTR(L, L->a, oe, SUP(L, eo, mu, s));
        sa_o[s] = hop_action(L->beta_a, ax, axp, ld_su2(UPTR(L, L->u, mu, eo, s)));
        sa_n[s] = hop_action(L->beta_a, ax, axp, h[s]);
      }
      for (int s = 0; s < n; ++s) { /* :167-174 */
        double rx = fed ? oc_draw(L) : oc_pu(L, pass, s, kp ? 6 : 0);
        if (exp(sa_o[s] - sa_n[s]) >= rx) {
          if (L->renorm_on_write) renorm_field(&h[s].a.re, 1);
          st_su2(UPTR(L, L->u, mu, eo, s), h[s]);
          acc1 += 1.0;
        }
      }
      if (measflag == 1) { /* :179-185 */
        double e = 0.0;
        for (int s = 0; s < n; ++s)
          e += m_nn(ld_su2(UPTR(L, L->u, mu, eo, s)), ld_su2(UPTR(L, L->stap, mu, eo, s))).a.re;
        L->plaq0 += e;
      }
    }
  }
  double vol = (double)L->v;
  if (kp) {
    *accept = acc_kp;
    if (measflag == 1) { /* :195-200 */
      L->plaq0 = L->plaq0 / (12.0 * vol);
      *accept = *accept / (3.0 * vol);
      acc1 = acc1 / (12.0 * vol);
      *accept = *accept * acc1;
    }
  } else {
    if (measflag == 1) L->plaq0 = L->plaq0 / (12.0 * vol);
    *accept = acc1 / (3.0 * vol);
  }
  free(b);
  free(a0o);
  free(sa_o);
  free(sa_n);
  free(vh);
  free(w);
  free(h);
}

void oc_agauge(oc_lattice *L, int measflag, double *accept) {
  link_update(L, 1, measflag, accept);
}
void oc_link_or(oc_lattice *L, int measflag, double *accept) {
  link_update(L, 0, measflag, accept);
}

/* -------------------------------------------------------------- a_force */

/* f/a_force.f:7-115: F(x) = sum_mu extract[U(x) X(a(x+mu)) U(x)^+ +
 *                                          U(x-mu)^+ X(a(x-mu)) U(x-mu)] */
void oc_a_force(oc_lattice *L, int eo) {
  int oe = 1 - eo, n = L->n;
  su2 *t1 = (su2 *)malloc(n * sizeof(su2)), *t2 = (su2 *)malloc(n * sizeof(su2));
  memset(L->force, 0, 3 * (size_t)n * sizeof(double));
  for (int mu = 0; mu < 3; ++mu) {
    for (int y = 0; y < n; ++y) { /* :64-68 */
      su2 uy = ld_su2(UPTR(L, L->u, mu, oe, y));
      t2[y] = m_dn(uy, m_an(APTR(L, L->a, oe, y), uy));
    }
    for (int x = 0; x < n; ++x) { /* :85-89 */
      su2 ux = ld_su2(UPTR(L, L->u, mu, eo, x));
      t1[x] = m_nn(ux, m_ad(APTR(L, L->a, oe, SUP(L, eo, mu, x)), ux));
    }
    for (int x = 0; x < n; ++x) { /* :99-108 */
      su2 q = t2[SDN(L, eo, mu, x)];
      double *f = L->force + 3 * (size_t)x;
      f[2] = f[2] + c_add(t1[x].a, q.a).im;
      f[1] = f[1] + c_add(t1[x].b, q.b).re;
      f[0] = f[0] + c_add(t1[x].b, q.b).im;
    }
  }
  free(t1);
  free(t2);
}

/* --------------------------------------------------------------- gauss */

/* f/gauss.f:10-31: X = sqrt(-ln(w1)/R) cos(2 pi w2), all w1 then all w2.
 * Philox mode: w1, w2 = draws kbase, kbase+1 of the site. */
static void gauss_core(oc_lattice *L, double *x, const double *r, int nh,
                       uint64_t pass, int kbase) {
  const double tpi = 2.0 * acos(-1.0);
  int fed = L->rng_mode == OC_RNG_FED;
  double *w1 = (double *)malloc(nh * sizeof(double));
  for (int i = 0; i < nh; ++i) w1[i] = fed ? oc_draw(L) : oc_pu(L, pass, i, kbase);
  for (int i = 0; i < nh; ++i) w1[i] = -1.0 * (1.0 / r[i]) * log(w1[i]);
  for (int i = 0; i < nh; ++i) {
    double w2 = fed ? oc_draw(L) : oc_pu(L, pass, i, kbase + 1);
    x[i] = sqrt(w1[i]) * cos(tpi * w2);
  }
  free(w1);
}

void oc_gauss(oc_lattice *L, double *x, const double *r, int nh) {
  uint64_t pass = L->pass;
  if (L->rng_mode == OC_RNG_PHILOX) L->pass += 1;
  gauss_core(L, x, r, nh, pass, 0);
}

/* -------------------------------------------------------- Higgs updates */

static void higgs_reset(oc_lattice *L) {
  L->a0kin = 0.0;
  L->a02 = 0.0;
  L->a04 = 0.0;
}

/* f/ahiggs1.f:7-127 (Metropolis with self-tuned step) */
void oc_ahiggs1(oc_lattice *L, double *delta, int measflag, double *accept) {
  int n = L->n, fed = L->rng_mode == OC_RNG_FED;
  double *an = (double *)malloc(3 * (size_t)n * sizeof(double));
  double *sfo = (double *)malloc(n * sizeof(double)), *sfn = (double *)malloc(n * sizeof(double));
  double *ro = (double *)malloc(n * sizeof(double)), *rn = (double *)malloc(n * sizeof(double));
  double acc = 0.0;
  higgs_reset(L);
  for (int eo = 0; eo < 2; ++eo) {
    uint64_t pass = L->pass;
    if (!fed) L->pass += 1;
    oc_a_force(L, eo);
    for (int s = 0; s < n; ++s) /* :51-56 */
      for (int i = 0; i < 3; ++i) {
        double rx = fed ? oc_draw(L) : oc_pu(L, pass, s, i);
        an[3 * s + i] = APTR(L, L->a, eo, s)[i] + (rx - 0.5) * (*delta);
      }
    for (int s = 0; s < n; ++s) { /* :61-73 */
      const double *ao = APTR(L, L->a, eo, s), *f = L->force + 3 * (size_t)s;
      sfo[s] = sfn[s] = ro[s] = rn[s] = 0.0;
      for (int i = 0; i < 3; ++i) {
        sfo[s] = sfo[s] + ao[i] * f[i];
        sfn[s] = sfn[s] + an[3 * s + i] * f[i];
        ro[s] = ro[s] + ao[i] * ao[i];
        rn[s] = rn[s] + an[3 * s + i] * an[3 * s + i];
      }
    }
    for (int s = 0; s < n; ++s) { /* :77-92 */
      double boltz = exp(-(L->beta_a * (sfo[s] - sfn[s])) + L->beta_2 * (ro[s] - rn[s]) +
                         L->beta_4 * (ro[s] * ro[s] - rn[s] * rn[s]));
      double rx = fed ? oc_draw(L) : oc_pu(L, pass, s, 3);
      if (boltz >= rx) {
        double *ao = APTR(L, L->a, eo, s);
        ao[0] = an[3 * s];
        ao[1] = an[3 * s + 1];
        ao[2] = an[3 * s + 2];
        ro[s] = rn[s];
        sfo[s] = sfn[s];
        acc += 1.0;
      }
    }
    if (measflag == 1) { /* :96-100 */
      double s2 = 0.0, s4 = 0.0, sk = 0.0;
      for (int s = 0; s < n; ++s) s2 += ro[s];
      for (int s = 0; s < n; ++s) s4 += ro[s] * ro[s];
      for (int s = 0; s < n; ++s) sk += sfo[s];
      L->a02 += s2;
      L->a04 += s4;
      L->a0kin += sk;
    }
  }
  double vol = (double)L->v;
  if (measflag == 1) { /* :107-111 */
    L->a02 /= vol;
    L->a04 /= vol;
    L->a0kin /= (6.0 * vol);
  }
  *accept = acc / vol; /* :113 */
  if (fabs(*accept - 0.5) > 0.1) /* :117-123 */
    *delta = (*accept > 0.5) ? 1.1131241 * (*delta) : 0.8989669 * (*delta);
  free(an);
  free(sfo);
  free(sfn);
  free(ro);
  free(rn);
}

/* f/ahiggs2.f (heatbath, or=0) and f/ahiggs3.f (overrelaxation, or=1) */
static void higgs_hb_or(oc_lattice *L, int orx, int measflag, double *accept) {
  int n = L->n, fed = L->rng_mode == OC_RNG_FED;
  double *alph = (double *)malloc(n * sizeof(double)), *width = (double *)malloc(n * sizeof(double));
  double *cx = (double *)malloc(n * sizeof(double));
  double *g = (double *)malloc(3 * (size_t)n * sizeof(double));
  double *an = (double *)malloc(3 * (size_t)n * sizeof(double));
  double *ro = (double *)malloc(n * sizeof(double)), *rn = (double *)malloc(n * sizeof(double));
  double acc = 0.0;
  higgs_reset(L);
  for (int eo = 0; eo < 2; ++eo) {
    uint64_t pass = L->pass;
    if (!fed) L->pass += 1;
    oc_a_force(L, eo);
    if (L->imode == 0) { /* f/ahiggs2.f:52-56 */
      for (int s = 0; s < n; ++s) {
        alph[s] = 0.0;
        width[s] = L->beta_2;
        cx[s] = 1.0 / L->beta_2 / 2.0;
      }
    }
    if (L->imode == 1) { /* f/ahiggs2.f:61-77 */
      double *conf2 = ro; /* scratch */
      for (int s = 0; s < n; ++s) conf2[s] = 0.0;
      for (int i = 0; i < 3; ++i)
        for (int s = 0; s < n; ++s) {
          double f = L->force[3 * (size_t)s + i];
          conf2[s] = conf2[s] + L->beta_a * (f * f);
        }
      for (int s = 0; s < n; ++s) {
        alph[s] = L->beta_4 * conf2[s];
        double d = L->beta_2 + L->beta_4 * conf2[s];
        alph[s] = alph[s] / (2.0 * (d * d));
        width[s] = L->beta_2 + alph[s];
        cx[s] = 1.0 / 2.0 / width[s];
      }
    }
    if (!orx) { /* f/ahiggs2.f:83-91 */
      for (int i = 0; i < 3; ++i) gauss_core(L, g + (size_t)i * n, width, n, pass, 2 * i);
      for (int s = 0; s < n; ++s)
        for (int i = 0; i < 3; ++i)
          an[3 * s + i] = g[(size_t)i * n + s] + L->beta_a * L->force[3 * (size_t)s + i] * cx[s];
    } else { /* f/ahiggs3.f:83-87 */
      for (int s = 0; s < n; ++s)
        for (int i = 0; i < 3; ++i)
          an[3 * s + i] = -APTR(L, L->a, eo, s)[i] +
                          2.0 * L->beta_a * L->force[3 * (size_t)s + i] * cx[s];
    }
    for (int s = 0; s < n; ++s) ro[s] = rn[s] = 0.0; /* f/ahiggs2.f:95-101 */
    for (int i = 0; i < 3; ++i)
      for (int s = 0; s < n; ++s) {
        double ao = APTR(L, L->a, eo, s)[i], a2 = an[3 * s + i];
        ro[s] = ro[s] + ao * ao;
        rn[s] = rn[s] + a2 * a2;
      }
    for (int s = 0; s < n; ++s) { /* f/ahiggs2.f:105-118 */
      double boltz = exp(-(alph[s] * (ro[s] - rn[s])) +
                         L->beta_4 * (ro[s] * ro[s] - rn[s] * rn[s]));
      double rx = fed ? oc_draw(L) : oc_pu(L, pass, s, orx ? 0 : 6);
      if (boltz >= rx) {
        double *ao = APTR(L, L->a, eo, s);
        ao[0] = an[3 * s];
        ao[1] = an[3 * s + 1];
        ao[2] = an[3 * s + 2];
        ro[s] = rn[s];
        acc += 1.0;
      }
    }
    if (measflag == 1) { /* f/ahiggs2.f:122-131 */
      double s2 = 0.0, s4 = 0.0;
      for (int s = 0; s < n; ++s) s2 += ro[s];
      for (int s = 0; s < n; ++s) s4 += ro[s] * ro[s];
      L->a02 += s2;
      L->a04 += s4;
      for (int i = 0; i < 3; ++i)
        for (int s = 0; s < n; ++s)
          L->a0kin = L->a0kin + APTR(L, L->a, eo, s)[i] * L->force[3 * (size_t)s + i];
    }
  }
  if (measflag == 1) { /* f/ahiggs2.f:138-143 */
    double vol = (double)L->v;
    L->a02 /= vol;
    L->a04 /= vol;
    L->a0kin /= (6.0 * vol);
    *accept = acc / vol;
  }
  free(alph);
  free(width);
  free(cx);
  free(g);
  free(an);
  free(ro);
  free(rn);
}

void oc_ahiggs2(oc_lattice *L, int measflag, double *accept) {
  higgs_hb_or(L, 0, measflag, accept);
}
void oc_ahiggs3(oc_lattice *L, int measflag, double *accept) {
  higgs_hb_or(L, 1, measflag, accept);
}

/* --------------------------------------------------------------- renorm */

static void renorm_field(double *u, size_t nlinks) {
  for (size_t k = 0; k < nlinks; ++k) {
    double *q = u + 4 * k;
    double t = (q[0] * q[0] + q[1] * q[1]) + (q[2] * q[2] + q[3] * q[3]);
    t = 1.0 / sqrt(t);
    for (int c = 0; c < 4; ++c) q[c] = t * q[c];
  }
}
/* f/renorm.f:6-37 */
void oc_renorm(oc_lattice *L) { renorm_field(L->u, 6 * (size_t)L->n); }
/* f/u_f_renorm.f:6-37 */
void oc_u_f_renorm(oc_lattice *L) { renorm_field(L->uf, 6 * (size_t)L->n); }

/* --------------------------------------------------------- gauge fixing */

/* f/or_gauge_fixing1.f:5-341 */
void oc_or_gauge_fixing(oc_lattice *L, double *da1) {
  int n = L->n;
  su2 *om = (su2 *)malloc(n * sizeof(su2));
  double da11 = 0.0;
  for (int eo = 0; eo < 2; ++eo) {
    int oe = 1 - eo;
    double dsum = 0.0;
    for (int s = 0; s < n; ++s) {
      su2 s1, s2;
      s1.a = c_mk(0.0, 0.0);
      s1.b = c_mk(0.0, 0.0);
      s2 = s1;
      for (int mu = 0; mu < 2; ++mu) { /* :97-102 */
        su2 u = ld_su2(UPTR(L, L->uf, mu, eo, s));
        s1.a = c_add(s1.a, u.a);
        s1.b = c_add(s1.b, u.b);
      }
      su2 u3 = ld_su2(UPTR(L, L->uf, 2, eo, s)); /* :104-107 */
      s1.a = c_add(s1.a, c_scale(L->alpha, u3.a));
      s1.b = c_add(s1.b, c_scale(L->alpha, u3.b));
      for (int mu = 0; mu < 2; ++mu) { /* :116-130 */
        su2 u = ld_su2(UPTR(L, L->uf, mu, oe, SDN(L, eo, mu, s)));
        s2.a = c_add(s2.a, u.a);
        s2.b = c_add(s2.b, u.b);
      }
      su2 d3 = ld_su2(UPTR(L, L->uf, 2, oe, SDN(L, eo, 2, s))); /* :139-142 */
      s2.a = c_add(s2.a, c_scale(L->alpha, d3.a));
      s2.b = c_add(s2.b, c_scale(L->alpha, d3.b));
      su2 w; /* :145-146 */
      w.a = c_add(s1.a, c_conj(s2.a));
      w.b = c_sub(s1.b, s2.b);
      if (eo == 1) { /* :153-158 */
        double d = c_sub(s1.a, s2.a).im;
        dsum += d * d + (w.b.re * w.b.re + w.b.im * w.b.im);
      }
      double k = (w.a.re * w.a.re + w.a.im * w.a.im) + (w.b.re * w.b.re + w.b.im * w.b.im);
      double sk = sqrt(k); /* :169-183 */
      su2 o;
      o.a = c_mk(w.a.re / sk, w.a.im / sk);
      o.b = c_mk(w.b.re / sk, w.b.im / sk);
      double ra = o.a.re;
      if (ra > 1.0) ra = 1.0;
      if (ra < -1.0) ra = -1.0;
      double th = acos(ra);
      if (th != 0.0) {
        double tho = L->lambda * th;
        double t1 = cos(tho);
        double t2 = sin(tho) / sin(th) * o.a.im;
        o.a = c_mk(t1, -t2);
        double f = sin(tho) / sin(th);
        o.b = c_mk(-(f * o.b.re), -(f * o.b.im));
      }
      om[s] = o;
    }
    if (eo == 1) da11 = dsum / (double)n;
    for (int s = 0; s < n; ++s) { /* :243-252 */
      double *af = APTR(L, L->af, eo, s);
      su2 t2 = m_nn(om[s], m_ad(af, om[s]));
      af[0] = t2.b.im;
      af[1] = t2.b.re;
      af[2] = t2.a.im;
    }
    for (int mu = 0; mu < 3; ++mu) { /* :256-324 */
      for (int s = 0; s < n; ++s)
        st_su2(UPTR(L, L->uf, mu, eo, s), m_nn(om[s], ld_su2(UPTR(L, L->uf, mu, eo, s))));
      for (int y = 0; y < n; ++y)
        st_su2(UPTR(L, L->uf, mu, oe, y),
               m_nd(ld_su2(UPTR(L, L->uf, mu, oe, y)), om[SUP(L, oe, mu, y)]));
    }
  }
  *da1 = da11; /* :331-337, one rank */
  free(om);
}

/* f/landau_gauge.f:6-63, with an iteration cap (the reference has none) */
void oc_landau_gauge(oc_lattice *L, double *da1, int *iter_or1) {
  memcpy(L->uf, L->u, 24 * (size_t)L->n * sizeof(double));
  memcpy(L->af, L->a, 6 * (size_t)L->n * sizeof(double));
  *da1 = 10000.0;
  *iter_or1 = 0;
  while (*da1 > L->limit_or) {
    if (*iter_or1 >= L->max_iter) {
      L->err = 2;
      return;
    }
    *iter_or1 += 1;
    oc_or_gauge_fixing(L, da1);
    oc_u_f_renorm(L);
  }
}

/* ----------------------------------------------------------- correlators */

static void corr_1d(const double *sl, int lz, double *corr) {
  for (int z = 0; z < lz; ++z) {
    double c = 0.0;
    for (int x0 = 0; x0 < lz; ++x0) c = c + sl[x0] * sl[(x0 + z) % lz];
    corr[z] = c / lz;
  }
}

/* f/correlator.f:11-186 (slice sums, p=0 projection, periodic z correlators)
 * and the a_f average of f/meas2a.f:39-48 */
void oc_correlator(oc_lattice *L, double *corr_a, double *corr_u, double *corr_3) {
  int lz = L->lz, h2 = L->h2;
  double *as = (double *)calloc(3 * (size_t)lz, sizeof(double));
  cpx *us = (cpx *)calloc(2 * 3 * (size_t)lz, sizeof(cpx)); /* [mu][slice][c] */
  for (int eo = 0; eo < 2; ++eo) { /* :53-79 */
    int start = 0;
    for (int sl = 0; sl < lz; ++sl) {
      for (int s = 0; s < h2; ++s) {
        const double *af = APTR(L, L->af, eo, start + s);
        for (int i = 0; i < 3; ++i) as[3 * sl + i] = as[3 * sl + i] + af[i];
        for (int mu = 0; mu < 3; ++mu) {
          const double *q = UPTR(L, L->uf, mu, eo, start + s);
          cpx *t = us + ((size_t)mu * lz + sl) * 2;
          t[0] = c_add(t[0], c_mk(q[1], 0.0));
          t[1] = c_add(t[1], c_mk(q[3], -q[2])); /* -i * beta */
        }
      }
      start += h2;
    }
  }
  double area = (double)(L->lx * L->ly); /* :97-100 */
  for (int k = 0; k < 3 * lz; ++k) as[k] = as[k] / area;
  for (int k = 0; k < 6 * lz; ++k) us[k] = c_mk(us[k].re / area, us[k].im / area);
  for (int z = 0; z < lz; ++z) { /* :139-158 */
    double ca = 0.0, cu = 0.0, c3 = 0.0;
    for (int x0 = 0; x0 < lz; ++x0) {
      int x1 = (x0 + z) % lz;
      double d = 0.0;
      for (int i = 0; i < 3; ++i) d = d + as[3 * x0 + i] * as[3 * x1 + i];
      ca = ca + d;
      cpx *p1 = us + ((size_t)0 * lz + x0) * 2, *q1 = us + ((size_t)0 * lz + x1) * 2;
      cpx *p2 = us + ((size_t)1 * lz + x0) * 2, *q2 = us + ((size_t)1 * lz + x1) * 2;
      cpx *p3 = us + ((size_t)2 * lz + x0) * 2, *q3 = us + ((size_t)2 * lz + x1) * 2;
      cpx t = c_add(c_add(c_add(c_mul(p1[0], q1[0]), c_mul(p1[1], c_conj(q1[1]))),
                          c_mul(p2[0], q2[0])),
                    c_mul(p2[1], c_conj(q2[1])));
      cu = cu + t.re;
      c3 = c3 + c_add(c_mul(p3[0], q3[0]), c_mul(p3[1], c_conj(q3[1]))).re;
    }
    corr_a[z] = ca;
    corr_u[z] = cu;
    corr_3[z] = c3;
  }
  for (int z = 0; z < lz; ++z) { /* :162-164 */
    corr_a[z] = corr_a[z] / lz;
    corr_u[z] = corr_u[z] / (lz * 2);
    corr_3[z] = corr_3[z] / lz;
  }
  for (int i = 0; i < 3; ++i) L->a0f[i] = 0.0; /* f/meas2a.f:39-48 */
  for (int eo = 0; eo < 2; ++eo)
    for (int s = 0; s < L->n; ++s)
      for (int i = 0; i < 3; ++i) L->a0f[i] = L->a0f[i] + APTR(L, L->af, eo, s)[i];
  for (int i = 0; i < 3; ++i) L->a0f[i] = L->a0f[i] / (double)L->v;
  free(as);
  free(us);
}

/* f/scalarcorr.f:10-119 on the unfixed a: slice means of rho = |a|^2 */
void oc_scalarcorr(oc_lattice *L, double *mean, double *corr) {
  int lz = L->lz, h2 = L->h2;
  double *sl = (double *)calloc(lz, sizeof(double));
  for (int eo = 0; eo < 2; ++eo) {
    int start = 0;
    for (int z = 0; z < lz; ++z) {
      for (int s = 0; s < h2; ++s) {
        const double *a = APTR(L, L->a, eo, start + s);
        sl[z] = sl[z] + a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
      }
      start += h2;
    }
  }
  double area = (double)(L->lx * L->ly), m = 0.0;
  for (int z = 0; z < lz; ++z) sl[z] = sl[z] / area;
  corr_1d(sl, lz, corr);
  for (int z = 0; z < lz; ++z) m += sl[z];
  *mean = m / lz;
  free(sl);
}

/* f/polyakov.f:10-123: slice means of P = cos(4 |a|), writes mean(P)^2 */
void oc_polyakov(oc_lattice *L, double *mean_sq, double *corr) {
  int lz = L->lz, h2 = L->h2;
  double *sl = (double *)calloc(lz, sizeof(double));
  for (int eo = 0; eo < 2; ++eo) {
    int start = 0;
    for (int z = 0; z < lz; ++z) {
      for (int s = 0; s < h2; ++s) {
        const double *a = APTR(L, L->a, eo, start + s);
        sl[z] = sl[z] + cos(4.0 * sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]));
      }
      start += h2;
    }
  }
  double area = (double)(L->lx * L->ly), m = 0.0;
  for (int z = 0; z < lz; ++z) sl[z] = sl[z] / area;
  corr_1d(sl, lz, corr);
  for (int z = 0; z < lz; ++z) m += sl[z];
  m = m / lz;
  *mean_sq = m * m;
  free(sl);
}

/* ---- learned Higgs proposal (liblceq's lceq_higgs_mlp; no reference
 * counterpart).  Action difference as f/ahiggs1.f:64-92. */
#define MLP_NIN 10
#define MLP_NH 64
#define MLP_NOUT 6
#define MLP_OFF_B1 (MLP_NIN * MLP_NH)
#define MLP_OFF_W2 (MLP_OFF_B1 + MLP_NH)
#define MLP_OFF_B2 (MLP_OFF_W2 + MLP_NH * MLP_NH)
#define MLP_OFF_W3 (MLP_OFF_B2 + MLP_NH)
#define MLP_OFF_B3 (MLP_OFF_W3 + MLP_NH * MLP_NOUT)
#define MLP_NPARAM (MLP_OFF_B3 + MLP_NOUT)

static float bf16_round(float f) { /* round to nearest even, as the GPU conversion */
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  u &= 0xFFFF0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

static void mlp_layer(const float *x, int nin, const float *w, const float *b, int nout, float *y,
                      int act) {
  for (int n = 0; n < nout; ++n) {
    double acc = 0.0;
    for (int k = 0; k < nin; ++k) acc += (double)x[k] * (double)bf16_round(w[k * nout + n]);
    float z = (float)acc + b[n];
    y[n] = act ? bf16_round(z / (1.0f + expf(-z))) : z;
  }
}

static void mlp_forward(const float *st, const float *raw, float *out) {
  float x[MLP_NIN], h1[MLP_NH], h2[MLP_NH];
  for (int k = 0; k < MLP_NIN; ++k)
    x[k] = bf16_round((raw[k] - st[MLP_NPARAM + k]) * st[MLP_NPARAM + MLP_NIN + k]);
  mlp_layer(x, MLP_NIN, st, st + MLP_OFF_B1, MLP_NH, h1, 1);
  mlp_layer(h1, MLP_NH, st + MLP_OFF_W2, st + MLP_OFF_B2, MLP_NH, h2, 1);
  mlp_layer(h2, MLP_NH, st + MLP_OFF_W3, st + MLP_OFF_B3, MLP_NOUT, out, 0);
}

/* frame with first axis F/|F|; second axis from z (or x when F is within
 * acos(0.9) of z) */
static void mlp_frame(const double F[3], double e[3][3], double *fn) {
  *fn = sqrt(F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  double e1[3] = {1.0, 0.0, 0.0};
  if (*fn > 0.0)
    for (int i = 0; i < 3; ++i) e1[i] = F[i] / *fn;
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
    e[0][i] = e1[i];
    e[1][i] = e2[i];
    e[2][i] = e3[i];
  }
}

static void mlp_raw(const double y[3], double fn, const oc_lattice *L, float *x) {
  x[0] = (float)y[0];
  x[1] = (float)y[1];
  x[2] = (float)y[2];
  x[3] = (float)fn;
  x[4] = 0.0f;
  x[5] = 0.0f;
  x[6] = (float)fn;
  x[7] = (float)L->beta_a;
  x[8] = (float)L->beta_2;
  x[9] = (float)L->beta_4;
}

static void mlp_proposal(const float *o, double mu[3], double sig[3], double ls[3]) {
  for (int i = 0; i < 3; ++i) {
    mu[i] = (double)o[i];
    ls[i] = (double)fminf(fmaxf(o[3 + i], -7.0f), 4.0f);
    sig[i] = exp(ls[i]);
  }
}

static double mlp_logq(const double y[3], const double mu[3], const double sig[3],
                       const double ls[3]) {
  double r = 0.0;
  for (int i = 0; i < 3; ++i) {
    double t = (y[i] - mu[i]) / sig[i];
    r = r - ls[i] - 0.5 * (t * t);
  }
  return r;
}

void oc_higgs_mlp(oc_lattice *L, const float *st, double *accept) {
  const double TPI = 2.0 * acos(-1.0);
  int n = L->n;
  double acc = 0.0;
  higgs_reset(L);
  for (int eo = 0; eo < 2; ++eo) {
    uint64_t pass = L->pass;
    L->pass += 1;
    oc_a_force(L, eo);
    double s2 = 0.0, s4 = 0.0, sk = 0.0;
    for (int s = 0; s < n; ++s) {
      double *a = APTR(L, L->a, eo, s);
      const double *F = L->force + 3 * (size_t)s;
      double e[3][3], fn, ya[3], yp[3], ypr[3], an[3];
      mlp_frame(F, e, &fn);
      for (int i = 0; i < 3; ++i) ya[i] = (e[i][0] * a[0] + e[i][1] * a[1]) + e[i][2] * a[2];
      float raw[MLP_NIN], o1[MLP_NOUT], o2[MLP_NOUT];
      mlp_raw(ya, fn, L, raw);
      mlp_forward(st, raw, o1);
      double mu[3], sig[3], ls[3], mu2[3], sig2[3], ls2[3];
      mlp_proposal(o1, mu, sig, ls);
      double u0 = oc_pu(L, pass, s, 0), u1 = oc_pu(L, pass, s, 1);
      double u2 = oc_pu(L, pass, s, 2), u3 = oc_pu(L, pass, s, 3), racc = oc_pu(L, pass, s, 4);
      double r01 = sqrt(-2.0 * log(u0)), r23 = sqrt(-2.0 * log(u2));
      double z[3] = {r01 * cos(TPI * u1), r01 * sin(TPI * u1), r23 * cos(TPI * u3)};
      for (int i = 0; i < 3; ++i) yp[i] = mu[i] + sig[i] * z[i];
      for (int i = 0; i < 3; ++i) an[i] = (e[0][i] * yp[0] + e[1][i] * yp[1]) + e[2][i] * yp[2];
      for (int i = 0; i < 3; ++i) ypr[i] = (e[i][0] * an[0] + e[i][1] * an[1]) + e[i][2] * an[2];
      mlp_raw(ypr, fn, L, raw);
      mlp_forward(st, raw, o2);
      mlp_proposal(o2, mu2, sig2, ls2);
      double lqf = mlp_logq(ypr, mu, sig, ls), lqr = mlp_logq(ya, mu2, sig2, ls2);
      double sfo = 0.0, sfn = 0.0, ro = 0.0, rn = 0.0;
      for (int i = 0; i < 3; ++i) {
        sfo = sfo + a[i] * F[i];
        sfn = sfn + an[i] * F[i];
        ro = ro + a[i] * a[i];
        rn = rn + an[i] * an[i];
      }
      double boltz = exp(-(L->beta_a * (sfo - sfn)) + L->beta_2 * (ro - rn) +
                         L->beta_4 * (ro * ro - rn * rn) + (lqr - lqf));
      if (boltz >= racc) {
        a[0] = an[0];
        a[1] = an[1];
        a[2] = an[2];
        ro = rn;
        acc += 1.0;
      }
      s2 += ro;
      s4 += ro * ro;
      sk += (a[0] * F[0] + a[1] * F[1]) + a[2] * F[2];
    }
    L->a02 += s2;
    L->a04 += s4;
    L->a0kin += sk;
  }
  double vol = (double)L->v;
  L->a02 /= vol;
  L->a04 /= vol;
  L->a0kin /= (6.0 * vol);
  *accept = acc / vol;
}

void oc_observables(const oc_lattice *L, double *out) {
  out[0] = L->plaq0;
  out[1] = L->a0kin;
  out[2] = L->a02;
  out[3] = L->a04;
  out[4] = L->a0f[0];
  out[5] = L->a0f[1];
  out[6] = L->a0f[2];
}
