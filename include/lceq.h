/*
 * lceq.h -- C ABI of the MI355X (gfx950) lattice update and measurement
 * library liblceq.so.  Every hot-path entry point replaces one Fortran
 * subroutine of derek-teaney/lceqcd; its comment cites the routine (file:line)
 * and the calls it replaces.  Fortran callers bind these with iso_c_binding
 * (lceqcd_amd/fortran/lceq_c.f90); the reference's implicit-interface
 * arguments map one to one (INTEGER -> int, REAL under -fdefault-real-8 ->
 * double).
 *
 * The lattice is device resident.  The reference keeps the fields in COMMON
 * /linkvar/ (f/INCLUDE/gaugefield.inc:71-86); here they move between host and
 * device only through lceq_put_* / lceq_get_*, which take arrays in that COMMON
 * layout with an explicit leading site dimension ld (the reference's maxarray,
 * or N = V/2):
 *   links  u(2 complex, ld, 0:1, 3)  -> double u[((mu*2 + p)*ld + s)*4 + c],
 *          c = Re alpha, Im alpha, Re beta, Im beta, mu = 0..2, p = 0 "even"
 *   higgs  a(3, ld, 0:1)             -> double a[(p*ld + s)*3 + i]
 * The staple v and the force follow the same conventions.
 *
 * All functions return LCEQ_OK (0) or a negative status; lceq_last_error()
 * describes the last failure on the calling thread.  One lattice handle must
 * not be used from two threads at once.  Work is queued on the handle's own
 * HIP stream; functions that return host values synchronise that stream.
 */
#ifndef LCEQ_H
#define LCEQ_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LCEQ_OK 0
#define LCEQ_ERR_ARG (-1)         /* invalid argument or lattice shape */
#define LCEQ_ERR_HIP (-2)         /* HIP runtime failure */
#define LCEQ_ERR_RNG (-3)         /* fed uniform stream exhausted */
#define LCEQ_ERR_NOT_CONVERGED (-4) /* landau_gauge hit max_iter */
#define LCEQ_ERR_IO (-5)          /* configuration file missing, short or unwritable */
#define LCEQ_ERR_FIELD (-6)       /* lceq_check_fields: non-finite or non-unitary fields */

/* uniform sources */
#define LCEQ_RNG_FED 0
#define LCEQ_RNG_PHILOX 1

/* sweep schedules for lceq_sweep */
#define LCEQ_SWEEP_KP_OR_METRO 0  /* 1 KP heatbath + n_or link OR + ahiggs1 */
#define LCEQ_SWEEP_REFERENCE 1    /* ahiggs2, ahiggs3, agauge, renorm (f/main.f:37-41) */

/* kernel classes for lceq_profile_begin */
#define LCEQ_KC_LINK_KP 1
#define LCEQ_KC_LINK_OR 2
#define LCEQ_KC_HIGGS 3
#define LCEQ_KC_GFIX 4

typedef struct lceq_lattice lceq_lattice;

/* ---- lifecycle -------------------------------------------------------- */
/* Whole lattice lx*ly*lz on the current HIP device; extents even and >= 2
 * (f/topology.f:124-142).  Replaces initialize/topology/adresses
 * (f/initialize.f:35-45, f/adresses.f:18) for one device. */
int lceq_create(int lx, int ly, int lz, lceq_lattice **out);
int lceq_destroy(lceq_lattice *L);
/* the extents lx, ly, lz of the lattice */
int lceq_dims(lceq_lattice *L, int *dims);
const char *lceq_last_error(void);
int lceq_sync(lceq_lattice *L);

/* ---- parameters (COMMON /paramet/, f/INCLUDE/input_parameter.inc:66-69) */
/* beta, beta_a, beta_2, beta_4 in the reference normalisation
 * (f/paraminit.f:150-153 derives them from beta, x, y). */
int lceq_set_couplings(lceq_lattice *L, double beta, double beta_a, double beta_2,
                       double beta_4);
/* lambda (overrelaxation), alpha (z weight), limit_or; max_iter caps
 * landau_gauge, which has no cap in the reference (f/landau_gauge.f:54-59). */
int lceq_set_gauge_fixing(lceq_lattice *L, double lambda, double alpha, double limit_or,
                          int max_iter);
/* imode of f/ahiggs2.f:52-77 (0 plain, 1 alpha-corrected width) */
int lceq_set_imode(lceq_lattice *L, int imode);
/* launch shape of the link and Higgs updates (no counterpart in the
 * reference): threads (= sites) per block 64, 128, 256 or 512, XCD-aware
 * block order (1) or launch order (0).  Fields do not depend on the shape;
 * plaquette and Higgs sums are added in a shape-dependent order.  Default
 * 256, XCD order. */
int lceq_set_shape(lceq_lattice *L, int threads_per_block, int xcd_order);
int lceq_get_shape(lceq_lattice *L, int *threads_per_block, int *xcd_order);
/* autotuner: time `trials` production sweeps (lceq_sweep schedule 0 with
 * n_or) at every shape and keep the fastest; fields, observables, the pass
 * counter and the history are restored.  ms (or NULL) receives 8 times per
 * sweep, shapes (64,128,256,512) x (launch order, XCD order).  Philox only. */
int lceq_autotune(lceq_lattice *L, int n_or, int trials, double *ms);
/* site order of the link and Higgs updates (no counterpart in the
 * reference): strip_rows = 0 takes the sites in index order; a power of two
 * >= 2 dividing ly (with ly * lz < 2^22) takes the rows (y, z) of a parity
 * in strips of that many y rows, each strip advancing through z, so that the
 * z neighbours of the sites in flight stay in the L2 of a lattice larger than
 * the Infinity Cache.
 * Fields do not depend on the order; plaquette and Higgs sums are added in an
 * order-dependent way.  Default 0. */
int lceq_set_site_order(lceq_lattice *L, int strip_rows);
int lceq_get_site_order(lceq_lattice *L, int *strip_rows);
/* time `trials` production sweeps with site orders 0, 4, 8, 16 (those that
 * divide ly) at the current launch shape, keep the fastest; fields,
 * observables, the pass counter and the history are restored.  ms (or NULL)
 * receives 4 times per sweep (-1 where an order does not fit). */
int lceq_autotune_site_order(lceq_lattice *L, int n_or, int trials, double *ms);
/* Measurement aid with no counterpart in the reference: a plain device copy
 * of a `bytes` buffer (a multiple of 16; one 16-byte element per thread),
 * 3 untimed copies, then `reps` copies timed with HIP events.  us_per_copy
 * receives the average; 2*bytes/us_per_copy is the practical read+write
 * bandwidth that the lattice launches are compared with.  Allocates and frees
 * two buffers of `bytes`. */
int lceq_copy_bandwidth(int64_t bytes, int reps, double *us_per_copy);
/* Measurement aid with no counterpart in the reference: `streams` (8, 16 or
 * 30) device arrays of `elements` 16-byte values read at the same index by
 * each thread, one 16-byte array written -- the access shape of a launch that
 * gathers that many operand arrays -- 3 untimed passes, then `reps` timed
 * with HIP events.  us_per_pass receives the average; (streams + 1) *
 * elements * 16 bytes move per pass.  Allocates and frees its buffers. */
int lceq_stream_bandwidth(int streams, int64_t elements, int reps, double *us_per_pass);

/* ---- uniforms --------------------------------------------------------- */
/* Counter-based Philox4x32-10 streams: draw k of site s in update pass P uses
 * counter (s, k/2, P_lo, P_hi) and key seed.  Each link (parity, mu) class and
 * each Higgs parity is one pass; pass_counter is the next pass number. */
int lceq_rng_philox(lceq_lattice *L, uint64_t seed, uint64_t pass_counter);
/* Fed mode: the routines consume `uniforms` in exactly the order the
 * reference's random_number calls do (f/heat_kpi.f:33-93, f/agauge.f:167-169,
 * f/gauss.f:21-26, f/ahiggs1.f:51-83, f/ahiggs2.f:83-111, f/ahiggs3.f:109-111). */
int lceq_rng_feed(lceq_lattice *L, const double *uniforms, int64_t n);
int lceq_rng_state(lceq_lattice *L, int64_t *fed_consumed, uint64_t *pass_counter);

/* ---- fields ----------------------------------------------------------- */
int lceq_put_links(lceq_lattice *L, const double *u, int64_t ld);
int lceq_get_links(lceq_lattice *L, double *u, int64_t ld);
int lceq_put_higgs(lceq_lattice *L, const double *a, int64_t ld);
int lceq_get_higgs(lceq_lattice *L, double *a, int64_t ld);
int lceq_put_links_fixed(lceq_lattice *L, const double *u_f, int64_t ld);
int lceq_get_links_fixed(lceq_lattice *L, double *u_f, int64_t ld);
int lceq_put_higgs_fixed(lceq_lattice *L, const double *a_f, int64_t ld);
int lceq_get_higgs_fixed(lceq_lattice *L, double *a_f, int64_t ld);
/* staple v of the last lceq_staples call (COMMON /helpfields/ v) */
int lceq_get_staple(lceq_lattice *L, double *v, int64_t ld);
/* force(3, ld) of the last lceq_a_force call (COMMON /helpfields/ force) */
int lceq_get_force(lceq_lattice *L, double *force, int64_t ld);
/* unitconf, f/confget.f:295-307 */
int lceq_unitconf(lceq_lattice *L);
/* testconf, f/confget.f:528-575 (one rank) */
int lceq_testconf(lceq_lattice *L);
/* hot start: Haar links and N(0, sigma^2) Higgs components (Philox stream) */
int lceq_hotstart(lceq_lattice *L, uint64_t seed, double sigma);

/* ---- hot path ----------------------------------------------------------- */
/* staples(eo,oe,mu), f/staples.f:9 (called at f/agauge.f:65); mu = 1..3 */
int lceq_staples(lceq_lattice *L, int eo, int mu);
/* heat_kpi(b,v,a0_old,nacch), f/heat_kpi.f:9 (called at f/agauge.f:93) on n
 * host sites; v gets 4 doubles (Re a, Im a, Re b, Im b) per site. */
int lceq_heat_kpi(lceq_lattice *L, int n, const double *b, double *v, const double *a0_old,
                  int *nacch);
/* agauge(measflag,accept), f/agauge.f:7 (called at f/main.f:23,40) */
int lceq_agauge(lceq_lattice *L, int measflag, double *accept);
/* link overrelaxation U -> Vhat^+ U^+ Vhat^+ with the hopping Metropolis of
 * f/agauge.f:115-174; no counterpart in the reference.  accept = Metropolis
 * acceptance (0..1). */
int lceq_link_or(lceq_lattice *L, int measflag, double *accept);
/* a_force(eo), f/a_force.f:7 (called at f/ahiggs1.f:47, f/ahiggs2.f:48,
 * f/ahiggs3.f:48) */
int lceq_a_force(lceq_lattice *L, int eo);
/* gauss(X,R,NH), f/gauss.f:10 (called at f/ahiggs2.f:83-85) on host arrays */
int lceq_gauss(lceq_lattice *L, double *x, const double *r, int nh);
/* ahiggs1(delta,measflag,accept), f/ahiggs1.f:7 */
int lceq_ahiggs1(lceq_lattice *L, double *delta, int measflag, double *accept);
/* ahiggs2(measflag,accept), f/ahiggs2.f:7 (called at f/main.f:21,38) */
int lceq_ahiggs2(lceq_lattice *L, int measflag, double *accept);
/* ahiggs3(measflag,accept), f/ahiggs3.f:7 (called at f/main.f:22,39) */
int lceq_ahiggs3(lceq_lattice *L, int measflag, double *accept);
/* renorm(), f/renorm.f:6 (called at f/main.f:24,41) */
int lceq_renorm(lceq_lattice *L);
/* Field guard (no reference counterpart; the reference has no checks and
 * aborts only on MPI errors): max_dev = max over links of
 * | |alpha|^2 + |beta|^2 - 1 | (finite links), nonfinite = number of NaN or
 * infinite link and Higgs values.  Returns LCEQ_ERR_FIELD when nonfinite > 0
 * or, for tol >= 0, when max_dev > tol.  Either output may be NULL. */
int lceq_check_fields(lceq_lattice *L, double tol, double *max_dev, int64_t *nonfinite);
/* lceq_check_fields without memory the library allocates.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under stream capture, serving loops, harnesses that forbid hipMalloc).  The
 * per-block partial sums and their one-block reduction live in a fixed 1 MiB
 * device array per GPU, held until the process ends, instead of the scratch
 * lceq_check_fields allocates on first use.  Lattices of at most 2^24 sites
 * per parity (2^16 blocks of 256 sites) are accepted; larger ones return
 * LCEQ_ERR_ARG before any device work.  Outputs and status equal
 * lceq_check_fields'.  Calls that overlap on one GPU (from two threads, or on
 * two lattices' streams) share the array and race. */
int lceq_check_fields_static(lceq_lattice *L, double tol, double *max_dev, int64_t *nonfinite);
/* u_f_renorm(), f/u_f_renorm.f:6 (called at f/landau_gauge.f:57) */
int lceq_u_f_renorm(lceq_lattice *L);
/* or_gauge_fixing(da1), f/or_gauge_fixing1.f:5 (called at f/landau_gauge.f:56) */
int lceq_or_gauge_fixing(lceq_lattice *L, double *da1);
/* landau_gauge(da1,iter_or1), f/landau_gauge.f:6 (called at f/main.f:62,74):
 * copies u->u_f, a->a_f and iterates or_gauge_fixing + u_f_renorm while
 * da1 > limit_or.  Returns LCEQ_ERR_NOT_CONVERGED if max_iter is reached. */
int lceq_landau_gauge(lceq_lattice *L, double *da1, int *iter_or1);
/* correlator(meas,mstart), f/correlator.f:11 (called at f/meas2a.f:35):
 * the three z-correlators the reference writes (arrays of length lz); also
 * sets a0f (f/meas2a.f:39-48) for lceq_observables. */
int lceq_correlator(lceq_lattice *L, double *corr_a, double *corr_u, double *corr_3);
/* scalarcorr(meas,mstart), f/scalarcorr.f:10 (called at f/main.f:70) */
int lceq_scalarcorr(lceq_lattice *L, double *mean, double *corr);
/* polyakov(meas,mstart), f/polyakov.f:10 (called at f/main.f:69,78) */
int lceq_polyakov(lceq_lattice *L, double *mean_sq, double *corr);
/* MULTALL{NN,DN,ND,DD,AD,AN}, f/matmul.f:1-194 (kind 0..5) on host arrays;
 * s1/s2 are 0-based site maps or NULL (normalsites). */
int lceq_multall(lceq_lattice *L, int kind, const double *op1, const double *op2, double *out,
                 int n, const int *s1, const int *s2);
/* COMMON /aver_loc/ (f/INCLUDE/measurement.inc:30): out[0..6] = plaq0,
 * a0kin, a02, a04, a0f(1..3) as set by the last routine that writes each. */
int lceq_observables(lceq_lattice *L, double *out);

/* ---- production sweeps (device resident, no host synchronisation) ------ */
/* nsweep sweeps of `schedule`; with measflag = 1 each sweep appends one row
 * of 16 doubles (plaq0, a0kin, a02, a04, a0f[3], link accept, Higgs accept,
 * da, iter, delta, ...) to the history.  delta (schedule 0) is the ahiggs1
 * step, tuned on device as f/ahiggs1.f:117-123 does; on return it holds the
 * current value.  The schedule-0 link writes are renormalised. */
int lceq_sweep(lceq_lattice *L, int schedule, int nsweep, int n_or, double *delta,
               int measflag);
/* copy the last n history rows (n <= rows recorded) and clear the history */
int lceq_history(lceq_lattice *L, double *out, int n);

/* ---- timing (HIP events on the lattice stream) -------------------------- */
/* time the launches of one kernel class until lceq_profile_end, which
 * returns their summed duration and count.  One event pair brackets each run
 * of consecutive launches of the class (e.g. the 6 * n_or overrelaxation
 * launches of a schedule-0 sweep), so no event sits between them; the
 * duration of a run includes the dispatch gaps inside it. */
int lceq_profile_begin(lceq_lattice *L, int kernel_class, int max_launches);
int lceq_profile_end(lceq_lattice *L, double *total_ms, int *launches);

/* ---- batched replicas ---------------------------------------------------- */
/* Many independent lattices of uneven sizes and couplings (a phase-diagram
 * scan, one Markov chain each) swept together: every (parity, mu) link
 * launch and every Higgs launch covers all replicas of a replica group (all
 * replicas, when the batch fits the Infinity Cache), and eight sweeps are one
 * replayed hipGraph.  No reference counterpart (the reference runs one
 * lattice per MPI job).  Replica r follows exactly the chain of a single
 * lattice with the same seed under lceq_sweep(LCEQ_SWEEP_KP_OR_METRO) (and,
 * with lceq_batch_sweep_mlp, under lceq_sweep_mlp): a device pass counter
 * numbers the Philox passes as the single lattice's pass counter does. */
typedef struct lceq_batch lceq_batch;
typedef struct lceq_mlp lceq_mlp;
/* dims[3r..3r+2] = extents of replica r (even, >= 2) */
int lceq_batch_create(int nrep, const int *dims, lceq_batch **out);
int lceq_batch_destroy(lceq_batch *B);
int lceq_batch_info(lceq_batch *B, int *nrep, int64_t *total_sites);
/* coup[4r..4r+3] = beta, beta_a, beta_2, beta_4 of replica r (reference normalisation) */
int lceq_batch_set_couplings(lceq_batch *B, const double *coup);
/* per-replica Philox keys; resets the sweep (pass) counter to 0 */
int lceq_batch_rng_philox(lceq_batch *B, const uint64_t *seeds);
int lceq_batch_hotstart(lceq_batch *B, const uint64_t *seeds, double sigma);
int lceq_batch_put_links(lceq_batch *B, int r, const double *u, int64_t ld);
int lceq_batch_get_links(lceq_batch *B, int r, double *u, int64_t ld);
int lceq_batch_put_higgs(lceq_batch *B, int r, const double *a, int64_t ld);
int lceq_batch_get_higgs(lceq_batch *B, int r, double *a, int64_t ld);
/* nsweep sweeps of 1 KP + n_or OR + ahiggs1 for every replica; delta (nrep
 * values, or NULL to keep the device values, initially 0.4) seeds the
 * ahiggs1 steps; measflag records one history row per replica per sweep;
 * use_graph replays a hipGraph of 8 captured sweeps (the remainder of nsweep
 * is launched directly).  Queued, no sync. */
int lceq_batch_sweep(lceq_batch *B, int nsweep, int n_or, const double *delta, int measflag,
                     int use_graph);
int lceq_batch_sync(lceq_batch *B);
/* the last n sweeps' rows, out[(k*nrep + r)*16 + i] (i as in lceq_sweep) */
int lceq_batch_history(lceq_batch *B, double *out, int n);
/* the current 16 observables of every replica, out[r*16 + i] */
int lceq_batch_observables(lceq_batch *B, double *out);
/* launch shape of the batch's update kernels (no counterpart in the
 * reference): threads (= sites) per block 64, 128, 256 or 512, and the
 * XCD-aware block order (1) or the launch order (0).  Fields do not depend on
 * the shape; the plaquette and Higgs sums are added in a shape-dependent
 * order.  Default 256, XCD order. */
int lceq_batch_set_shape(lceq_batch *B, int threads_per_block, int xcd_order);
int lceq_batch_get_shape(lceq_batch *B, int *threads_per_block, int *xcd_order);
/* replica groups (no counterpart in the reference): a sweep runs its
 * launches over one run of consecutive replicas at a time, so that a batch
 * larger than the 256 MB Infinity Cache is swept in groups whose fields stay
 * in it from launch to launch.  The replicas are split into
 * ceil(fields / group_bytes) runs of about equal field bytes (120 bytes per
 * site); 0 sweeps the batch as one group.  The chains do not depend on the
 * grouping.  Default 288 MiB. */
int lceq_batch_set_group_bytes(lceq_batch *B, int64_t group_bytes);
/* group_bytes, the number of groups, and the first replica of each of the
 * first max_groups groups (any pointer may be NULL) */
int lceq_batch_get_groups(lceq_batch *B, int64_t *group_bytes, int *ngroups, int *first_replica,
                          int max_groups);
/* autotuner: time `trials` sweeps of every shape on this batch, keep the
 * fastest; fields, observables and the sweep counter are restored.  ms (or
 * NULL) receives 8 times per sweep, shapes (64,128,256,512) x (launch order,
 * XCD order). */
int lceq_batch_autotune(lceq_batch *B, int n_or, int trials, double *ms);
/* nsweep sweeps of 1 KP + n_or OR + the learned Higgs proposal of M for
 * every replica, one network for all replicas (each replica's couplings are
 * network inputs); history rows as lceq_batch_sweep, the Higgs observables
 * always measured (as lceq_sweep_mlp).  The learned launches take 256-site
 * blocks whatever the launch shape.  No reference counterpart; the plain
 * update it replaces is f/ahiggs1.f:43-123.  Queued, no sync. */
int lceq_batch_sweep_mlp(lceq_batch *B, lceq_mlp *M, int nsweep, int n_or, int measflag,
                         int use_graph);
/* training records of M from every replica at its own couplings, as
 * lceq_mlp_collect: per parity the inputs at the current F, nsteps ahiggs1
 * steps of that parity for every replica (kept in the chains; one Philox pass
 * each, no history row), the new a in the same frame; every stride-th site */
int lceq_batch_mlp_collect(lceq_batch *B, lceq_mlp *M, int nsteps, int stride);

/* ---- learned Higgs proposal (no counterpart in the reference) -------------
 * Per site, a 10 -> 64 -> 64 -> 6 SiLU network (bf16 MFMA inference, fp32
 * accumulation) maps the local a and the hopping force F of f/a_force.f,
 * both in the frame whose first axis is F/|F|, |F| and the couplings to the
 * mean and log-scale of a Gaussian proposal in that frame.  The proposal is
 * accepted with the FP64 Metropolis-Hastings probability min(1, exp(S(a) -
 * S(a')) q(a|a')/q(a'|a)), S the site action of f/ahiggs1.f, so the chain is
 * exact for any weights; the plain Metropolis update (lceq_ahiggs1) is its
 * fallback.  Training is maximum likelihood on moves of the plain update. */
#define LCEQ_MLP_NPARAM 5254 /* W1[10][64] b1[64] W2[64][64] b2[64] W3[64][6] b3[6] */
#define LCEQ_MLP_NSTATE 5274 /* parameters, feature mean[10], feature 1/scale[10] */
/* Glorot-uniform weights from seed, zero biases, identity normalisation */
int lceq_mlp_create(uint64_t seed, lceq_mlp **out);
int lceq_mlp_destroy(lceq_mlp *M);
int lceq_mlp_get_state(lceq_mlp *M, float *state, int n);
int lceq_mlp_set_state(lceq_mlp *M, const float *state, int n);
/* training records: for each parity of the lattice's current configuration,
 * the inputs at the current F, nsteps plain Metropolis steps of that parity
 * (kept in the chain), the new a in the same frame; every stride-th site */
int lceq_mlp_collect(lceq_lattice *L, lceq_mlp *M, int nsteps, int stride);
int lceq_mlp_records(lceq_mlp *M, int64_t *n);
int lceq_mlp_clear(lceq_mlp *M);
/* steps Adam steps on mini-batches of `batch` records (a multiple of 32);
 * fit_norm first sets the input normalisation from the records; loss (or
 * NULL) receives the mean negative log-likelihood of every step */
int lceq_mlp_train(lceq_mlp *M, int steps, int batch, double lr, uint64_t seed, int fit_norm,
                   double *loss);
/* launch shape of the learned Higgs update: sites per wave 64 (four MFMA
 * column tiles, the default, also selected by 0) or 32.  Fields do not
 * depend on it. */
int lceq_mlp_set_sites_per_wave(lceq_mlp *M, int sites);
/* the inference network on n rows of the 10 raw inputs: out[n][6] */
int lceq_mlp_eval(lceq_mlp *M, const float *x, int n, float *out);
/* one Higgs sweep (both parities) with the learned proposal; accept is the
 * acceptance; the Higgs observables (rho, rho^2, a.F) are always measured */
int lceq_higgs_mlp(lceq_lattice *L, lceq_mlp *M, double *accept);
/* nsweep sweeps of 1 KP + n_or overrelaxation + the learned Higgs update,
 * history rows as lceq_sweep with measflag */
int lceq_sweep_mlp(lceq_lattice *L, lceq_mlp *M, int nsweep, int n_or, int measflag);

/* ---- configuration files (host) -------------------------------------------
 * The reference's direct-access unformatted configuration files
 * (f/confsave.f:50-253, f/confget.f:318-526): link records per (mu, parity,
 * z) of 16*2*Lx*Ly/2 bytes, Higgs records per (parity, z) of 8*3*Lx*Ly/2
 * bytes, each the xy half-plane in site order x/2 + (Lx/2) y.  real_bytes = 8
 * for a build with 8-byte reals (records full), 4 for the as-shipped build
 * (values in the first half of each record, the rest zero on write).  Either
 * path may be NULL to skip that field.  Host arrays use the lceq_put_links /
 * lceq_put_higgs layout with ld = N. */
int lceq_conf_write(const int *dims, const double *u, const double *a, const char *u_path,
                    const char *a_path, int real_bytes);
int lceq_conf_read(const int *dims, double *u, double *a, const char *u_path,
                   const char *a_path, int real_bytes);
/* CONFSAVE (f/confsave.f:18): download the lattice and write the files */
int lceq_confsave(lceq_lattice *L, const char *u_path, const char *a_path, int real_bytes);
/* OLDCONF (f/confget.f:318): read the files and upload them to the lattice */
int lceq_oldconf(lceq_lattice *L, const char *u_path, const char *a_path, int real_bytes);
/* The file name f/confsave.f:249-253 writes (upper = 1: '_BE', '_U') or the
 * one f/measinit.f:246-250 makes oldconf open (upper = 0: '_be', '_u'): the
 * two differ in case.  higgs selects a_conf_ over u_conf_; beta, x, y and
 * sweep as the reference passes them (nint(1000 beta), nint(10^4 x),
 * nint(10^4 y), confstart + nsweep) with Fortran I edit descriptors (a value
 * too wide for its field becomes asterisks, e.g. a negative y). */
int lceq_conf_filename(int higgs, int upper, int lz, double beta, double x, double y, int sweep,
                       char *out, int cap);

/* ---- RANLUX (host) --------------------------------------------------------
 * Luscher's luxury subtract-with-borrow generator, single precision, with the
 * state, restart vectors and luxury levels of the reference's seeding
 * generator (f/luxury.f).  Pure host code: no device, no lattice. */
typedef struct lceq_ranlux_state lceq_ranlux_state;
/* a generator that initialises itself with the defaults (luxury 3, seed
 * 314159265) at its first lceq_ranlux call, as f/luxury.f:31-66 */
int lceq_ranlux_create(lceq_ranlux_state **out);
int lceq_ranlux_destroy(lceq_ranlux_state *R);
/* RANLUX(RVEC, LENV), f/luxury.f:12-112: lenv uniforms in (0,1) */
int lceq_ranlux(lceq_ranlux_state *R, float *rvec, int lenv);
/* RLUXGO(LUX, INT, K1, K2), f/luxury.f:209-303: luxury 0..4 or a p of 24..2000
 * (lux < 0: default 3), seed ins (0: default), restart after k1 + 10^9 k2
 * numbers.  LCEQ_ERR_ARG where the reference prints an error (luxury out of
 * range, negative seed, impossible restart point), after the same fallbacks. */
int lceq_rluxgo(lceq_ranlux_state *R, int lux, int ins, int k1, int k2);
/* RLUXUT(ISDEXT), f/luxury.f:174-190: the 25-integer restart vector */
int lceq_rluxut(lceq_ranlux_state *R, int *isdext);
/* RLUXIN(ISDEXT), f/luxury.f:120-171: restore from a restart vector */
int lceq_rluxin(lceq_ranlux_state *R, const int *isdext);
/* RLUXAT(LOUT, INOUT, K1, K2), f/luxury.f:194-205 (NULL outputs are skipped) */
int lceq_rluxat(lceq_ranlux_state *R, int *lout, int *inout, int *k1, int *k2);
/* the 16 integer seeds MPI rank `rank` derives in f/measinit.f:214-216:
 * RLUXGO(3, 314159 + rank, 0, 0), 16 RANLUX numbers, iseed = int(10^6 u) */
int lceq_measinit_seeds(int rank, int *iseed);
/* the device streams of L seeded as f/measinit.f:214-217 seeds the
 * reference's random_number for rank `rank` (random_seed(put = iseed)): the
 * Philox key is lceq_measinit_key(iseed) of the rank's 16 measinit seeds and
 * the pass counter restarts at 0; iseed (or NULL) receives the 16 seeds */
int lceq_rng_measinit(lceq_lattice *L, int rank, int *iseed);
/* the Philox key of 16 measinit seeds: iseed[0] + 2^32 iseed[1] (each < 10^6) */
uint64_t lceq_measinit_key(const int *iseed);

#ifdef __cplusplus
}
#endif
#endif
