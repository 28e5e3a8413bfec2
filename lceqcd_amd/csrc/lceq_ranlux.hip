// lceq_ranlux.hip -- host-side RANLUX (Luscher's luxury version of the
// Marsaglia-Zaman subtract-with-borrow generator) with the state, restart
// vectors and luxury levels of f/luxury.f:12-303, and the per-rank seeding of
// f/measinit.f:214-217.  The reference uses it only to seed its uniform
// stream; here it derives seeds with the reference's recipe.  Arithmetic is
// single precision as in the reference build: every state value is a
// multiple of 2^-24 in [0,1), so the subtractions are exact.
//
// The algorithm is the published one: G. Marsaglia and A. Zaman, Ann. Appl.
// Probab. 1 (1991) 462 (subtract-with-borrow, lags 24 and 10, base 2^24);
// M. Luscher, Comput. Phys. Commun. 79 (1994) 100 (luxury levels: of every
// 24 + skip numbers only 24 are delivered, p = 24, 48, 97, 223, 389); F. James,
// Comput. Phys. Commun. 79 (1994) 111 (the portable implementation f/luxury.f
// follows: seeding by a multiplicative congruential generator, the 25-word
// restart vector, the initialisation by (level, seed, k1, k2)).  Its output
// bits and restart vectors are an interface (seeds in the reference's files),
// so the lags, carry, skip counts and counters keep that implementation's
// names and steps.
#include <stdint.h>
#include <string.h>

#include "../../include/lceq.h"
#include "lceq_internal.hpp"

namespace {
constexpr int MAXLEV = 4, LXDFLT = 3;
constexpr int JSDFLT = 314159265;
constexpr int IGIGA = 1000000000;
constexpr int ITWO24 = 1 << 24, ICONS = 2147483563;
constexpr int NDSKIP[MAXLEV + 1] = {0, 24, 73, 199, 365};  // p = 24, 48, 97, 223, 389
}  // namespace

struct lceq_ranlux_state {
  bool notyet;
  int i24, j24;  // 1-based lags, f/luxury.h COMMON /LUX1/
  float carry, seeds[25], twom24, twom12;
  int luxlev, nskip, in24, next[25], kount, mkount, inseed;
};

static float subtract_with_borrow(lceq_ranlux_state* R) {
  float uni = R->seeds[R->j24] - R->seeds[R->i24] - R->carry;
  if (uni < 0.0f) {
    uni = uni + 1.0f;
    R->carry = R->twom24;
  } else {
    R->carry = 0.0f;
  }
  R->seeds[R->i24] = uni;
  R->i24 = R->next[R->i24];
  R->j24 = R->next[R->j24];
  return uni;
}

// 24 seeds from the multiplicative congruential generator of
// f/luxury.f:50-62 (rluxgo: f/luxury.f:262-276)
static void seed_from_int(lceq_ranlux_state* R, int jseed) {
  R->twom24 = 1.0f;
  int iseeds[25];
  for (int i = 1; i <= 24; ++i) {
    R->twom24 = R->twom24 * 0.5f;
    int k = jseed / 53668;
    jseed = 40014 * (jseed - k * 53668) - k * 12211;
    if (jseed < 0) jseed = jseed + ICONS;
    iseeds[i] = jseed % ITWO24;
  }
  R->twom12 = R->twom24 * 4096.0f;
  for (int i = 1; i <= 24; ++i) {
    R->seeds[i] = (float)iseeds[i] * R->twom24;
    R->next[i] = i - 1;
  }
  R->next[1] = 24;
  R->i24 = 24;
  R->j24 = 10;
  R->carry = 0.0f;
  if (R->seeds[24] == 0.0f) R->carry = R->twom24;
}

extern "C" int lceq_ranlux_create(lceq_ranlux_state** out) {
  if (!out) return lceq::set_error(LCEQ_ERR_ARG, "null output");
  lceq_ranlux_state* R = new lceq_ranlux_state();
  memset(R, 0, sizeof(*R));
  R->notyet = true;
  R->i24 = 24;
  R->j24 = 10;
  R->luxlev = LXDFLT;
  *out = R;
  return LCEQ_OK;
}

extern "C" int lceq_ranlux_destroy(lceq_ranlux_state* R) {
  delete R;
  return LCEQ_OK;
}

// ranlux(rvec, lenv), f/luxury.f:12-112
extern "C" int lceq_ranlux(lceq_ranlux_state* R, float* rvec, int lenv) {
  if (!R || lenv < 0 || (lenv > 0 && !rvec)) return lceq::set_error(LCEQ_ERR_ARG, "bad ranlux arguments");
  if (R->notyet) {  // default initialisation, f/luxury.f:31-66
    R->notyet = false;
    R->inseed = JSDFLT;
    R->luxlev = LXDFLT;
    R->nskip = NDSKIP[R->luxlev];
    R->in24 = 0;
    R->kount = 0;
    R->mkount = 0;
    seed_from_int(R, JSDFLT);
  }
  for (int ivec = 0; ivec < lenv; ++ivec) {
    float uni = subtract_with_borrow(R);
    rvec[ivec] = uni;
    if (uni < R->twom12) {  // pad numbers with fewer than 12 significant bits
      rvec[ivec] = rvec[ivec] + R->twom24 * R->seeds[R->j24];
      if (rvec[ivec] == 0.0f) rvec[ivec] = R->twom24 * R->twom24;
    }
    if (++R->in24 == 24) {  // skip to the luxury level
      R->in24 = 0;
      R->kount = R->kount + R->nskip;
      for (int isk = 0; isk < R->nskip; ++isk) (void)subtract_with_borrow(R);
    }
  }
  R->kount = R->kount + lenv;
  if (R->kount >= IGIGA) {
    R->mkount = R->mkount + 1;
    R->kount = R->kount - IGIGA;
  }
  return LCEQ_OK;
}

// rluxut(isdext), f/luxury.f:174-190
extern "C" int lceq_rluxut(lceq_ranlux_state* R, int* isdext) {
  if (!R || !isdext) return lceq::set_error(LCEQ_ERR_ARG, "bad rluxut arguments");
  for (int i = 1; i <= 24; ++i) isdext[i - 1] = (int)(R->seeds[i] * 4096.0f * 4096.0f);
  isdext[24] = R->i24 + 100 * R->j24 + 10000 * R->in24 + 1000000 * R->luxlev;
  if (R->carry > 0.0f) isdext[24] = -isdext[24];
  return LCEQ_OK;
}

// rluxin(isdext), f/luxury.f:120-171
extern "C" int lceq_rluxin(lceq_ranlux_state* R, const int* isdext) {
  if (!R || !isdext) return lceq::set_error(LCEQ_ERR_ARG, "bad rluxin arguments");
  R->notyet = false;
  R->twom24 = 1.0f;
  for (int i = 1; i <= 24; ++i) {
    R->next[i] = i - 1;
    R->twom24 = R->twom24 * 0.5f;
  }
  R->next[1] = 24;
  R->twom12 = R->twom24 * 4096.0f;
  for (int i = 1; i <= 24; ++i) R->seeds[i] = (float)isdext[i - 1] * R->twom24;
  R->carry = 0.0f;
  if (isdext[24] < 0) R->carry = R->twom24;
  int isd = isdext[24] < 0 ? -isdext[24] : isdext[24];
  R->i24 = isd % 100;
  isd /= 100;
  R->j24 = isd % 100;
  isd /= 100;
  R->in24 = isd % 100;
  isd /= 100;
  R->luxlev = isd;
  if (R->luxlev <= MAXLEV) {
    R->nskip = NDSKIP[R->luxlev];
  } else if (R->luxlev >= 24) {
    R->nskip = R->luxlev - 24;
  } else {
    R->nskip = NDSKIP[MAXLEV];
    R->luxlev = MAXLEV;
  }
  R->inseed = -1;
  return LCEQ_OK;
}

// rluxat(lout, inout, k1, k2), f/luxury.f:194-205
extern "C" int lceq_rluxat(lceq_ranlux_state* R, int* lout, int* inout, int* k1, int* k2) {
  if (!R) return lceq::set_error(LCEQ_ERR_ARG, "null generator");
  if (lout) *lout = R->luxlev;
  if (inout) *inout = R->inseed;
  if (k1) *k1 = R->kount;
  if (k2) *k2 = R->mkount;
  return LCEQ_OK;
}

// rluxgo(lux, ins, k1, k2), f/luxury.f:209-303.  Returns LCEQ_ERR_ARG for the
// inputs the reference reports as errors (negative seed, restart point that
// cannot occur at this luxury level) after applying the same fallbacks.
extern "C" int lceq_rluxgo(lceq_ranlux_state* R, int lux, int ins, int k1, int k2) {
  if (!R) return lceq::set_error(LCEQ_ERR_ARG, "null generator");
  int rc = LCEQ_OK;
  if (lux < 0) {
    R->luxlev = LXDFLT;
  } else if (lux <= MAXLEV) {
    R->luxlev = lux;
  } else if (lux < 24 || lux > 2000) {
    R->luxlev = MAXLEV;
    rc = lceq::set_error(LCEQ_ERR_ARG, "illegal RANLUX luxury %d", lux);
  } else {
    R->luxlev = lux;
    for (int ilx = 0; ilx <= MAXLEV; ++ilx)
      if (lux == NDSKIP[ilx] + 24) R->luxlev = ilx;
  }
  R->nskip = R->luxlev <= MAXLEV ? NDSKIP[R->luxlev] : R->luxlev - 24;
  R->in24 = 0;
  if (ins < 0) rc = lceq::set_error(LCEQ_ERR_ARG, "illegal RANLUX seed %d", ins);
  int jseed = ins > 0 ? ins : JSDFLT;
  R->inseed = jseed;
  R->notyet = false;
  seed_from_int(R, jseed);
  R->kount = k1;
  R->mkount = k2;
  if (k1 + k2 != 0) {  // restart at a break point: skip k1 + IGIGA*k2
    for (int iouter = 1; iouter <= k2 + 1; ++iouter) {
      int inner = iouter == k2 + 1 ? k1 : IGIGA;
      for (int isk = 0; isk < inner; ++isk) (void)subtract_with_borrow(R);
    }
    R->in24 = R->kount % (R->nskip + 24);
    if (R->mkount > 0) {
      int izip = IGIGA % (R->nskip + 24);
      int izip2 = R->mkount * izip + R->in24;
      R->in24 = izip2 % (R->nskip + 24);
    }
    if (R->in24 > 23) {
      R->in24 = 0;
      rc = lceq::set_error(LCEQ_ERR_ARG, "RANLUX restart %d %d %d cannot occur at luxury %d", ins,
                           k1, k2, R->luxlev);
    }
  }
  return rc;
}

// the nseed = 16 integer seeds rank `rank` gets from f/measinit.f:214-216:
// rluxgo(3, 314159+rank, 0, 0); ranlux(rvec, 16); iseed = int(10**6 * rvec)
extern "C" int lceq_measinit_seeds(int rank, int* iseed) {
  if (!iseed || rank < 0) return lceq::set_error(LCEQ_ERR_ARG, "bad arguments");
  lceq_ranlux_state R;
  memset(&R, 0, sizeof(R));
  lceq_rluxgo(&R, 3, 314159 + rank, 0, 0);
  float rvec[16];
  lceq_ranlux(&R, rvec, 16);
  for (int i = 0; i < 16; ++i) iseed[i] = (int)(1000000.0f * rvec[i]);
  return LCEQ_OK;
}

// The device streams of lattice L seeded from rank `rank`'s measinit seeds,
// as f/measinit.f:214-217 seeds the reference's random_number with them
// (random_seed(put = iseed)): the Philox key is iseed(1) + 2^32 iseed(2)
// (each below 10^6), the pass counter restarts at 0.  iseed (or NULL)
// receives the 16 seeds, which the reference leaves in COMMON /paramet/.
extern "C" int lceq_rng_measinit(lceq_lattice* L, int rank, int* iseed) {
  if (!L) return lceq::set_error(LCEQ_ERR_ARG, "null lattice handle");
  int s[16];
  int rc = lceq_measinit_seeds(rank, s);
  if (rc) return rc;
  rc = lceq_rng_philox(L, lceq_measinit_key(s), 0);
  if (rc) return rc;
  if (iseed) memcpy(iseed, s, sizeof s);
  return LCEQ_OK;
}

extern "C" uint64_t lceq_measinit_key(const int* iseed) {
  if (!iseed) return 0;
  return (uint64_t)(uint32_t)iseed[0] | ((uint64_t)(uint32_t)iseed[1] << 32);
}
