// lceq_internal.hpp -- the lattice context shared by the C-ABI layer
// (lceq_api.hip) and the kernel launchers (lceq_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lceq_device.hpp"

// number of per-block partial quantities a reduction slot holds
#define LCEQ_NQ 4
// reduction slots per routine call (6 link sub-passes, 2 Higgs parities, ...)
#define LCEQ_NSLOT 8

// device observables (double obs[LCEQ_OBS_N])
enum {
  OBS_PLAQ0 = 0,   // plaq0 of the last link update (f/agauge.f:179-200)
  OBS_A0KIN = 1,   // a0kin of the last Higgs update (f/ahiggs2.f:126-147)
  OBS_A02 = 2,
  OBS_A04 = 3,
  OBS_A0F = 4,     // 4..6: a0f(1..3) (f/meas2a.f:39-48)
  OBS_ACC_LINK = 7,
  OBS_ACC_HIGGS = 8,
  OBS_DA = 9,
  OBS_ITER = 10,
  OBS_DELTA = 11,  // ahiggs1 Metropolis step (device-resident in sweeps)
  OBS_NACCH = 12,  // raw counts of the last link call
  OBS_NACC1 = 13,
  OBS_GF_DONE = 14,  // landau_gauge convergence flag (1.0 = converged)
  OBS_GF_CAPPED = 15,
  LCEQ_OBS_N = 16
};

struct lceq_lattice {
  lceq::Geom g;
  int vol;
  int device;
  hipStream_t stream;
  // device fields
  double *U, *A, *Uf, *Af, *Vst, *F;
  double *partials;  // [LCEQ_NSLOT][LCEQ_NQ][pstride]
  int pstride;       // >= number of 64-site blocks (the smallest launch shape)
  double *obs;       // [LCEQ_OBS_N]
  double *hist;      // [hist_cap][LCEQ_OBS_N] per-sweep records
  int hist_cap;
  int hist_rows;     // rows recorded since the last lceq_history
  double *slices;    // correlator slice partial sums [14][lz][corr_split]
  double *corr;      // correlator outputs [5][lz] + 4 scalars
  double *scratch;   // generic device scratch for host-array entry points
  size_t scratch_bytes;
  double *fed;       // fed uniforms
  int64_t fed_cap, fed_n, fed_pos;
  // parameters (reference COMMON /paramet/, f/INCLUDE/input_parameter.inc:66-69)
  double beta, beta_a, beta_2, beta_4, lambda, alpha, limit_or;
  int max_iter, imode;
  // RNG
  int rng_mode;  // 0 fed, 1 philox
  // launch shape of the link and Higgs updates: threads (= sites) per block
  // 64/128/256/512 and the XCD-aware block order (partials sized for 64)
  int tb, xcd;
  // site order of those launches: strips of `strip` y rows (0: index order)
  int strip;
  uint64_t seed, pass;
  // timing of one kernel class (bench instrumentation): one event pair per
  // run of consecutive launches of that class, launches counted inside runs
  int prof_class;
  int prof_n, prof_cap, prof_depth, prof_launches;
  hipEvent_t *prof_ev;
};

namespace lceq {

// record a failure for lceq_last_error() and return `code`
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Why an lx x ly x lz lattice cannot be created, or nullptr if it can.
// Besides even extents >= 2: 3 V links fit the 32-bit link counters, the
// index divisions need ly, lz < 2^22, and the device loads address a site
// of a class array as an unsigned 32-bit byte offset s * 16 (16-byte link
// halves) and the stores' raw buffer range is 2^32 - 1 bytes, so a parity
// holds fewer than 2^28 sites (s * 16 + 16 <= 2^32 - 16).  Checked before
// any device call.
inline const char* geom_reject(int lx, int ly, int lz) {
  if (lx < 2 || ly < 2 || lz < 2 || (lx & 1) || (ly & 1) || (lz & 1))
    return "extents must be even and >= 2";
  if ((int64_t)lx * ly * lz * 3 > INT32_MAX) return "lattice too large for 32-bit link counters";
  if (ly >= (1 << 22) || lz >= (1 << 22)) return "ly and lz must be below 2^22";
  if ((int64_t)lx * ly * lz / 2 >= ((int64_t)1 << 28))
    return "2^28 or more sites per parity (32-bit byte offsets of 16-byte link halves)";
  return nullptr;
}

enum KernelClass { KC_NONE = 0, KC_LINK_KP = 1, KC_LINK_OR = 2, KC_HIGGS = 3, KC_GFIX = 4 };

int blocks_for(int n);
// blocks of the link and Higgs update launches at the lattice's launch shape
int shape_blocks(const lceq_lattice* L);

// Event brackets around runs of launches of the profiled class.  Brackets
// nest (a sweep brackets its launches, each launcher brackets itself); only
// the outermost records events, so back-to-back launches of one class are
// timed without events between them and the average per launch matches the
// kernel's own duration up to the dispatch gaps inside the run.
inline void prof_begin(lceq_lattice* L, int kc) {
  if (L->prof_class != kc) return;
  if (L->prof_depth++ == 0 && L->prof_n + 2 <= L->prof_cap)
    (void)hipEventRecord(L->prof_ev[L->prof_n++], L->stream);
}
inline void prof_end(lceq_lattice* L, int kc) {
  if (L->prof_class != kc) return;
  if (--L->prof_depth == 0 && (L->prof_n & 1))
    (void)hipEventRecord(L->prof_ev[L->prof_n++], L->stream);
}
// one launch of class kc inside an open run
inline void prof_count(lceq_lattice* L, int kc) {
  if (L->prof_class == kc && (L->prof_n & 1)) ++L->prof_launches;
}

// launchers (lceq_kernels.hip); all enqueue on L->stream, no host sync.
// Link launches use reduction slots 0-5 ((eo, mu) order), Higgs launches
// slots 6-7 (eo); with reduce = 0 a launch writes no partial sums.
void launch_link(lceq_lattice* L, int kp, int eo, int mu, int measflag, int renorm_on_write,
                 const double* fed, uint64_t pass, int reduce);
// One-block reduction of the partial sums into the observables: link_kind
// 0 none, 1 KP (agauge), 2 overrelaxation; higgs_kind 0 none, 1-3 ahiggs1-3;
// with hist_row the observables are also copied into that history row.
// higgs_nb: block partials of the Higgs slots when their launch used another
// shape than the lattice's (0: the lattice's shape)
void launch_finalize(lceq_lattice* L, int link_kind, int link_meas, int higgs_kind,
                     int higgs_meas, double* hist_row, int higgs_nb = 0);
void launch_staples(lceq_lattice* L, int eo, int mu);
void launch_force(lceq_lattice* L, int eo);
void launch_higgs(lceq_lattice* L, int kind, int eo, int measflag, const double* fed,
                  uint64_t pass);
void launch_renorm(lceq_lattice* L, double* U);
// per-block (unitarity deviation, non-finite count) of U and A into part[2 nb]
void launch_check_fields(lceq_lattice* L, double* part);
// One half-iteration of gauge fixing; check = 1 first tests the loop
// condition for the `iter` iterations already complete (see k_gfix).
void launch_gfix(lceq_lattice* L, int eo, int compute_da, int renorm, int guarded, int check = 0,
                 int iter = 0);
void launch_gfix_finalize(lceq_lattice* L, int guarded, int iter = 0);
// blocks per z slice of the slice sums (the slices buffer holds 14 lz split)
int corr_split(const Geom& g);
void launch_slices(lceq_lattice* L);
void launch_corr(lceq_lattice* L);
void launch_heat_kpi(lceq_lattice* L, int n, const double* b, const double* a0_old, double* v,
                     const double* fed, uint64_t pass);
void launch_gauss(lceq_lattice* L, int n, const double* r, double* x, const double* fed,
                  uint64_t pass);
void launch_multall(lceq_lattice* L, int kind, const double* op1, const double* op2,
                    double* out, int n, const int* s1, const int* s2);
void launch_hotstart(lceq_lattice* L, uint64_t seed, double sigma);
void launch_unitconf(lceq_lattice* L);
void launch_testconf(lceq_lattice* L);

}  // namespace lceq

// helpers of the C ABI layer (lceq_api.hip) shared with the other entry files
int check_launch();
uint64_t next_pass(lceq_lattice* L);
int read_obs(lceq_lattice* L, double* host);
int write_obs(lceq_lattice* L, int index, double value);
// one full link sweep (six (eo, mu) launches) of the KP heatbath or the
// overrelaxation, partials in slots 0-5 when reduce is set
int link_sweep(lceq_lattice* L, int kp, int measflag, int renorm_on_write, int reduce);
