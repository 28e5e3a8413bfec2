/*
 * lceq_oracle.h -- CPU restatement of the lceqcd lattice update and
 * measurement path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (liblceq.so) never calls it.
 *
 * Field layout follows the reference COMMON blocks
 * (f/INCLUDE/gaugefield.inc:71-76) with the leading dimension equal to the
 * number of sites per parity N (no MPI halo tail on one rank):
 *   u[((mu*2 + p)*N + s)*4 + c]   c = Re alpha, Im alpha, Re beta, Im beta
 *   a[(p*N + s)*3 + i]            i = a(1), a(2), a(3)
 * mu = 0,1,2 is the reference's mu = 1,2,3; p = 0 is the reference's "even".
 */
#ifndef LCEQ_ORACLE_H
#define LCEQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_lattice oc_lattice;

/* uniform sources */
#define OC_RNG_FED 0
#define OC_RNG_PHILOX 1

oc_lattice *oc_create(int lx, int ly, int lz);
void oc_destroy(oc_lattice *L);
int oc_error(const oc_lattice *L);

void oc_set_couplings(oc_lattice *L, double beta, double beta_a, double beta_2,
                      double beta_4);
void oc_set_gauge_fixing(oc_lattice *L, double lambda, double alpha,
                         double limit_or, int max_iter);
void oc_set_imode(oc_lattice *L, int imode);
/* production sweeps (lceq_sweep schedule 0) normalise every accepted link
 * update as f/renorm.f does; 1 switches the link updates to that mode */
void oc_set_renorm_on_write(oc_lattice *L, int flag);

void oc_rng_feed(oc_lattice *L, const double *buf, int64_t n);
int64_t oc_rng_consumed(const oc_lattice *L);
void oc_rng_philox(oc_lattice *L, uint64_t seed, uint64_t pass_counter);
uint64_t oc_pass_counter(const oc_lattice *L);

/* geometry (restatement of f/adresses.f), 0-based, tab[(p*3+mu)*N + s] */
void oc_site_tables(const oc_lattice *L, int32_t *siteup, int32_t *sitedn);
void oc_site_tables_closed_form(const oc_lattice *L, int32_t *siteup,
                                int32_t *sitedn);

/* fields */
double *oc_links(oc_lattice *L);
double *oc_higgs(oc_lattice *L);
double *oc_links_fixed(oc_lattice *L);
double *oc_higgs_fixed(oc_lattice *L);
double *oc_staple(oc_lattice *L); /* v, same layout as u */
double *oc_force(oc_lattice *L);  /* force(3, N) */
void oc_unitconf(oc_lattice *L);
void oc_testconf(oc_lattice *L);
void oc_hotstart(oc_lattice *L, uint64_t seed, double sigma);

/* SU(2) algebra on site vectors (f/matmul.f); kind 0..5 = NN DN ND DD AD AN.
 * For AD/AN, op1 holds Higgs triples (3 doubles per entry). */
void oc_multall(int kind, const double *op1, const double *op2, double *out,
                int n, const int32_t *s1, const int32_t *s2);

/* hot path */
void oc_staples(oc_lattice *L, int eo, int mu);
void oc_heat_kpi(oc_lattice *L, int n, const double *b, double *v,
                 const double *a0_old, int64_t *nacch);
void oc_agauge(oc_lattice *L, int measflag, double *accept);
void oc_link_or(oc_lattice *L, int measflag, double *accept);
void oc_a_force(oc_lattice *L, int eo);
void oc_gauss(oc_lattice *L, double *x, const double *r, int nh);
void oc_ahiggs1(oc_lattice *L, double *delta, int measflag, double *accept);
void oc_ahiggs2(oc_lattice *L, int measflag, double *accept);
void oc_ahiggs3(oc_lattice *L, int measflag, double *accept);
void oc_renorm(oc_lattice *L);
void oc_u_f_renorm(oc_lattice *L);
void oc_or_gauge_fixing(oc_lattice *L, double *da1);
void oc_landau_gauge(oc_lattice *L, double *da1, int *iter_or1);
void oc_correlator(oc_lattice *L, double *corr_a, double *corr_u,
                   double *corr_3);
void oc_scalarcorr(oc_lattice *L, double *mean, double *corr);
void oc_polyakov(oc_lattice *L, double *mean_sq, double *corr);

/* observables written by the last update call (f/INCLUDE/measurement.inc:30)
 * out[0..6] = plaq0, a0kin, a02, a04, a0f(1..3) */
void oc_observables(const oc_lattice *L, double *out);

/* The learned Higgs proposal of liblceq (no reference counterpart): both
 * parities, network state = 5254 parameters + feature mean[10] + 1/scale[10]
 * (the library's layout).  The network is evaluated as the GPU's bf16
 * inference does (bf16 inputs, weights and hidden activations, products
 * summed in double and rounded to float, fp32 bias and SiLU); the frame, the
 * Box-Muller draws and the FP64 Metropolis-Hastings test are those of the
 * library.  Philox mode only.  Measures like ahiggs2 with measflag = 1. */
void oc_higgs_mlp(oc_lattice *L, const float *state, double *accept);

/* Philox4x32-10 block (Salmon et al., SC'11), for known-answer tests */
void oc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                      uint32_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
