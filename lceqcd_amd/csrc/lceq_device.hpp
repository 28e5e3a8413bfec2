// lceq_device.hpp -- device building blocks for the gfx950 lattice kernels:
// SU(2) algebra in (alpha, beta) form, the checkerboard geometry, the
// Philox4x32-10 uniform streams, and deterministic wavefront/LDS reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lceq {

// 1/sqrt(t) and cos, sin(2 pi r) in the update kernels.  The product library
// takes the reciprocal square root (within an ulp of 1/sqrt, one operation)
// and cospi / sincospi of 2r (no argument reduction by a rounded 2 pi; the
// reference's cos(tpi * r) differs by an ulp or so).  The test-only library
// built with LCEQ_ORACLE_FORMS (and without contraction) rounds them as the
// CPU restatement writes them, 1 / sqrt(t) and cos(tpi * r) with
// tpi = 2 acos(-1), so that its chains differ from the restatement's only
// where a libm function (log, cos, sin, exp) rounds differently.
__device__ __forceinline__ double inv_sqrt(double t) {
#ifdef LCEQ_ORACLE_FORMS
  return 1.0 / sqrt(t);
#else
  return rsqrt(t);
#endif
}
__device__ __forceinline__ double cos_2pi(double r) {
#ifdef LCEQ_ORACLE_FORMS
  return cos((2.0 * M_PI) * r);
#else
  return cospi(2.0 * r);
#endif
}
__device__ __forceinline__ void sincos_2pi(double r, double* sn, double* cs) {
#ifdef LCEQ_ORACLE_FORMS
  *sn = sin((2.0 * M_PI) * r);
  *cs = cos((2.0 * M_PI) * r);
#else
  sincospi(2.0 * r, sn, cs);
#endif
}

// ---------------------------------------------------------------- complex

struct cpx {
  double re, im;
};
__device__ __forceinline__ cpx c_mk(double r, double i) { return cpx{r, i}; }
__device__ __forceinline__ cpx c_add(cpx a, cpx b) { return cpx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cpx c_sub(cpx a, cpx b) { return cpx{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cpx c_neg(cpx a) { return cpx{-a.re, -a.im}; }
__device__ __forceinline__ cpx c_conj(cpx a) { return cpx{a.re, -a.im}; }
__device__ __forceinline__ cpx c_mul(cpx a, cpx b) {
  return cpx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cpx c_scale(double t, cpx a) { return cpx{t * a.re, t * a.im}; }

// U = [[alpha, beta], [-beta*, alpha*]]; the same form holds for sums of
// SU(2) matrices (staples) and for the algebra element X(a) = i a.sigma,
// (alpha, beta) = (i a3, a2 + i a1)  (f/INCLUDE/gaugefield.inc:89-101).
struct su2 {
  cpx a, b;
};

// Products in the operand conventions of f/matmul.f, written in the real
// components (Re a, Im a, Re b, Im b): each output is a sum of four products,
// one multiply and three fused multiply-adds (16 operations per product; the
// complex form costs 20).  DP / DQ take the adjoint of p / q, which flips the
// signs of Im a, Re b and Im b.
template <bool DP, bool DQ>
__device__ __forceinline__ su2 q_prod(su2 p, su2 q) {
  double p0 = p.a.re, p1 = DP ? -p.a.im : p.a.im, p2 = DP ? -p.b.re : p.b.re,
         p3 = DP ? -p.b.im : p.b.im;
  double q0 = q.a.re, q1 = DQ ? -q.a.im : q.a.im, q2 = DQ ? -q.b.re : q.b.re,
         q3 = DQ ? -q.b.im : q.b.im;
  return su2{c_mk(((p0 * q0 - p1 * q1) - p2 * q2) - p3 * q3,
                  ((p0 * q1 + p1 * q0) + p2 * q3) - p3 * q2),
             c_mk(((p0 * q2 - p1 * q3) + p2 * q0) + p3 * q1,
                  ((p0 * q3 + p1 * q2) - p2 * q1) + p3 * q0)};
}
// acc + p q, the chains started at acc (no separate additions)
template <bool DP, bool DQ>
__device__ __forceinline__ su2 q_prod_acc(su2 acc, su2 p, su2 q) {
  double p0 = p.a.re, p1 = DP ? -p.a.im : p.a.im, p2 = DP ? -p.b.re : p.b.re,
         p3 = DP ? -p.b.im : p.b.im;
  double q0 = q.a.re, q1 = DQ ? -q.a.im : q.a.im, q2 = DQ ? -q.b.re : q.b.re,
         q3 = DQ ? -q.b.im : q.b.im;
  return su2{c_mk((((acc.a.re + p0 * q0) - p1 * q1) - p2 * q2) - p3 * q3,
                  (((acc.a.im + p0 * q1) + p1 * q0) + p2 * q3) - p3 * q2),
             c_mk((((acc.b.re + p0 * q2) - p1 * q3) + p2 * q0) + p3 * q1,
                  (((acc.b.im + p0 * q3) + p1 * q2) - p2 * q1) + p3 * q0)};
}
__device__ __forceinline__ su2 m_nn(su2 p, su2 q) {  // U1 U2   (matmul.f:22-23)
  return q_prod<false, false>(p, q);
}
__device__ __forceinline__ su2 m_dn(su2 p, su2 q) {  // U1^+ U2 (matmul.f:52-56)
  return q_prod<true, false>(p, q);
}
__device__ __forceinline__ su2 m_nd(su2 p, su2 q) {  // U1 U2^+ (matmul.f:85-89)
  return q_prod<false, true>(p, q);
}
__device__ __forceinline__ su2 m_dd(su2 p, su2 q) {  // U1^+ U2^+ (matmul.f:118-122)
  return q_prod<true, true>(p, q);
}
struct vec3 {
  double x[3];
};
// (i y) z for real y: the product with x1 = (0, y) without its zero terms
// (0 * z.re is +-0 and adding it changes no finite value, so the result is
// the full complex product's, up to the sign of a zero)
__device__ __forceinline__ cpx c_mul_i(double y, cpx z) { return cpx{-(y * z.im), y * z.re}; }
__device__ __forceinline__ su2 m_ad(const vec3& a, su2 u) {  // X(a) U^+ (matmul.f:150-154)
  double y = a.x[2];
  cpx x2 = c_mk(a.x[1], a.x[0]);
  return su2{c_add(c_mul_i(y, c_conj(u.a)), c_mul(x2, c_conj(u.b))),
             c_add(c_neg(c_mul_i(y, u.b)), c_mul(x2, u.a))};
}
__device__ __forceinline__ su2 m_an(const vec3& a, su2 u) {  // X(a) U (matmul.f:183-187)
  double y = a.x[2];
  cpx x2 = c_mk(a.x[1], a.x[0]);
  return su2{c_sub(c_mul_i(y, u.a), c_mul(x2, c_conj(u.b))),
             c_add(c_mul_i(y, u.b), c_mul(x2, c_conj(u.a)))};
}

// U (b.sigma) U^+ = (R b).sigma for U = c0 + i c.sigma, i.e. alpha = c0 + i c3,
// beta = c2 + i c1:  R b = (c0^2 - |c|^2) b + 2 (c.b) c - 2 c0 (c x b).  In
// exact arithmetic this is the adjoint triple of the products U X(b) U^+
// (m_nn(U, m_ad(b, U)), f/a_force.f:54-110) for any U, unitary or not, at 22
// instead of about 34 operations.  TR: U^+ X(b) U, the transposed rotation.
template <bool TR>
__device__ __forceinline__ vec3 adj_rot(su2 u, const vec3& b) {
  double c0 = u.a.re, c1 = u.b.im, c2 = u.b.re, c3 = u.a.im;
  double s = c0 * c0 - ((c1 * c1 + c2 * c2) + c3 * c3);
  double k = 2.0 * ((c1 * b.x[0] + c2 * b.x[1]) + c3 * b.x[2]);
  double w = TR ? 2.0 * c0 : -2.0 * c0;
  double x0 = c2 * b.x[2] - c3 * b.x[1];  // c x b
  double x1 = c3 * b.x[0] - c1 * b.x[2];
  double x2 = c1 * b.x[1] - c2 * b.x[0];
  return vec3{{(s * b.x[0] + k * c1) + w * x0, (s * b.x[1] + k * c2) + w * x1,
               (s * b.x[2] + k * c3) + w * x2}};
}

// ------------------------------------------------------------ geometry

// Checkerboard numbering of f/adresses.f on one rank: a site (x,y,z), 0-based,
// has parity p = (x+y+z)&1 (p = 0 is the reference's "even") and index
// s = x/2 + h1*y + h2*z within its parity, h1 = lx/2, h2 = lx*ly/2.
struct Geom {
  int lx, ly, lz;
  int h1, h2, n;  // n = sites per parity
  int npad;       // leading dimension of every SoA component array
  float rh1, rh2;  // 1/h1, 1/h2 (geom_init) for the index divisions
  int wy, wz;      // (ly-1)*h1, (lz-1)*h2: the index steps that wrap y and z
};

// Geometry of an lx x ly x lz lattice (all even)
__host__ __device__ inline Geom geom_init(int lx, int ly, int lz) {
  Geom g;
  g.lx = lx;
  g.ly = ly;
  g.lz = lz;
  g.h1 = lx / 2;
  g.h2 = lx * ly / 2;
  g.n = lx * ly * lz / 2;
  g.npad = (g.n + 63) & ~63;
  g.rh1 = 1.0f / (float)g.h1;
  g.rh2 = 1.0f / (float)g.h2;
  g.wy = (ly - 1) * g.h1;
  g.wz = (lz - 1) * g.h2;
  return g;
}

// floor(s / d) for s < 2^31 and a quotient below 2^22: the single-precision
// estimate s * (1/d) is within one of the quotient, and one correction step
// either way makes it exact (no integer division sequence)
__device__ __forceinline__ int div_small(int s, int d, float rd) {
  int q = (int)((float)s * rd);
  int r = s - q * d;
  q += r >= d ? 1 : 0;
  q -= r < 0 ? 1 : 0;
  return q;
}

struct Coord {
  int x, y, z;
};

// Site order of an update launch.  sy = 0: the site index order (position P
// = site s).  sy = 2^k > 0: the rows (y, z) of the parity in strips of sy
// consecutive y, each strip advancing through z before the next strip; P
// runs along a row (h1 sites), then over the sy rows of one z, then z.  The
// rows of one z in a strip are contiguous in memory, and the z neighbours of
// a position are sy * h1 positions away instead of h2, so the blocks of one
// XCD reuse them from its L2 on lattices too large for the Infinity Cache.
struct SiteOrder {
  int sy, lsy;  // strip height and its log2
  int rows;     // sy * lz: rows per strip
  float rrows;  // 1 / rows
};
__host__ __device__ inline SiteOrder site_order(const Geom& g, int sy) {
  SiteOrder o;
  o.sy = sy;
  o.lsy = 0;
  while ((1 << o.lsy) < sy) ++o.lsy;
  o.rows = sy * g.lz;
  o.rrows = sy ? 1.0f / (float)o.rows : 0.0f;
  return o;
}

__device__ __forceinline__ Coord site_coord(const Geom& g, int p, int s) {
  int z = div_small(s, g.h2, g.rh2);
  int r = s - z * g.h2;
  int y = div_small(r, g.h1, g.rh1);
  int xh = r - y * g.h1;
  return Coord{2 * xh + ((p + y + z) & 1), y, z};
}
// the site at position P of the order o (P < N)
__device__ __forceinline__ unsigned order_site(const Geom& g, const SiteOrder& o, unsigned P) {
  if (!o.sy) return P;
  int r = div_small((int)P, g.h1, g.rh1);  // row in the order
  int k = (int)P - r * g.h1;
  int strip = div_small(r, o.rows, o.rrows);
  int q = r - strip * o.rows;
  int z = q >> o.lsy;
  int y = (strip << o.lsy) + (q - (z << o.lsy));
  return (unsigned)(k + g.h1 * y + g.h2 * z);
}

// Index changes of the moves +1 (p) and -1 (m) along x, y, z from a site at
// c, periodic: a y or z move keeps x/2 and changes the index by +-h1 or
// +-h2 (wrapping at the faces); an x move changes x/2 by 0 or +-1 with the
// parity of x (wrapping at x = 0, lx - 1).  Moves along different axes add,
// so the index of c + e_mu - e_nu is s + p[mu] + m[nu].
struct Steps {
  int p[3], m[3];
};
__device__ __forceinline__ Steps site_steps(const Geom& g, Coord c) {
  Steps t;
  bool odd = c.x & 1;
  t.p[0] = odd ? (c.x == g.lx - 1 ? 1 - g.h1 : 1) : 0;
  t.m[0] = odd ? 0 : (c.x == 0 ? g.h1 - 1 : -1);
  t.p[1] = c.y == g.ly - 1 ? -g.wy : g.h1;
  t.m[1] = c.y == 0 ? g.wy : -g.h1;
  t.p[2] = c.z == g.lz - 1 ? -g.wz : g.h2;
  t.m[2] = c.z == 0 ? g.wz : -g.h2;
  return t;
}
// component d (0..2, a run-time value) of a step vector
__device__ __forceinline__ int step_of(const int (&v)[3], int d) {
  return d == 0 ? v[0] : (d == 1 ? v[1] : v[2]);
}
__device__ __forceinline__ void unit(int mu, int& dx, int& dy, int& dz) {
  dx = mu == 0;
  dy = mu == 1;
  dz = mu == 2;
}

// ------------------------------------------------------------- storage

// Links: U_mu at parity p, site s as two 16-byte halves
//   ((double2*)U)[((mu*2 + p)*2 + h)*npad + s],
//   h = 0: (Re alpha, Im alpha), h = 1: (Re beta, Im beta),
// so a (mu, p) class still starts at U + (mu*2 + p)*4*npad.
// Higgs of parity p at A + p*3*npad: (a1, a2) as one double2 array of npad
// entries, then a3 as a double array of npad entries.
// A site costs two 16-byte loads per link and a 16- plus an 8-byte load per
// Higgs triple: every wave-wide load instruction costs the same texture
// addressing and data-return cycles whether a lane reads 8 or 16 bytes, so
// the wide halves halve that work per byte.  Class bases are uniform per
// launch and the site offset is an unsigned 32-bit byte offset, so the loads
// use the SGPR-base + 32-bit VGPR-offset form (below 4 GiB per array).
template <typename T>
__device__ __forceinline__ T& at(T* base, unsigned s) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + s * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ const T& at(const T* base, unsigned s) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + s * (unsigned)sizeof(T));
}
__device__ __forceinline__ su2 ld_link(const double* __restrict__ U, const Geom& g, int mu,
                                       int p, unsigned s) {
  const double2* q = reinterpret_cast<const double2*>(U + (size_t)((mu * 2 + p) * 4) * (size_t)g.npad);
  double2 lo = at(q, s), hi = at(q + g.npad, s);
  return su2{c_mk(lo.x, lo.y), c_mk(hi.x, hi.y)};
}
// Stores write through the L2 (device-scope cache policy sc1): every launch
// ends with the L2s of the eight XCDs written back, so that the next,
// dependent launch sees the updated fields on every XCD; stored through as
// they are made, the updated links are already out when the last wave ends.
// Raw buffer stores take the policy bits; the range check (offset + size <=
// 2^32 - 1) holds because a parity has fewer than 2^28 sites.
typedef int lceq_v4i __attribute__((ext_vector_type(4)));
typedef int lceq_v2i __attribute__((ext_vector_type(2)));
constexpr int CPOL_WRITE_THROUGH = 16;  // sc1
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(base, 0, -1, 0x00020000);
}
__device__ __forceinline__ void st_link(double* __restrict__ U, const Geom& g, int mu, int p,
                                        unsigned s, su2 u) {
  double2* q = reinterpret_cast<double2*>(U + (size_t)((mu * 2 + p) * 4) * (size_t)g.npad);
  __builtin_amdgcn_raw_buffer_store_b128(
      __builtin_bit_cast(lceq_v4i, make_double2(u.a.re, u.a.im)), raw_rsrc(q), s * 16u, 0,
      CPOL_WRITE_THROUGH);
  __builtin_amdgcn_raw_buffer_store_b128(
      __builtin_bit_cast(lceq_v4i, make_double2(u.b.re, u.b.im)), raw_rsrc(q + g.npad), s * 16u,
      0, CPOL_WRITE_THROUGH);
}
__device__ __forceinline__ vec3 ld_higgs(const double* __restrict__ A, const Geom& g, int p,
                                         unsigned s) {
  const double* q = A + (size_t)(p * 3) * (size_t)g.npad;
  double2 v = at(reinterpret_cast<const double2*>(q), s);
  return vec3{{v.x, v.y, at(q + 2 * (size_t)g.npad, s)}};
}
__device__ __forceinline__ void st_higgs(double* __restrict__ A, const Geom& g, int p, unsigned s,
                                         const vec3& a) {
  double* q = A + (size_t)(p * 3) * (size_t)g.npad;
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(lceq_v4i, make_double2(a.x[0], a.x[1])),
                                         raw_rsrc(q), s * 16u, 0, CPOL_WRITE_THROUGH);
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(lceq_v2i, a.x[2]),
                                        raw_rsrc(q + 2 * (size_t)g.npad), s * 8u, 0,
                                        CPOL_WRITE_THROUGH);
}

// Offset in doubles of component c (0..ncomp-1) of site s in class cls of a
// link (ncomp = 4) or Higgs (ncomp = 3) field in the layout above.
__host__ __device__ inline size_t field_offset(int ncomp, int cls, int c, size_t s, size_t npad) {
  size_t base = (size_t)cls * ncomp * npad;
  if (ncomp == 4) return base + (size_t)(c >> 1) * 2 * npad + 2 * s + (c & 1);
  return base + (c < 2 ? 2 * s + c : 2 * npad + s);
}

// ----------------------------------------------------------------- RNG

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Draw k of lattice
// index idx in pass `pass` uses counter (idx, (stream<<8) | k/2, pass_lo,
// pass_hi) and key = seed; words 2(k%2), 2(k%2)+1 give a 52-bit x and the
// uniform (2x+1) 2^-53 in (0,1).
struct Philox {
  uint32_t k0, k1;
};

__device__ __forceinline__ void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    // one 32x32->64 multiply (v_mad_u64_u32) per product instead of a
    // separate low and high half: both are quarter-rate instructions
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}
__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
  uint64_t x = ((uint64_t)hi << 20) | (uint64_t)(lo >> 12);
  return (double)(2 * x + 1) * 0x1p-53;
}
// the two uniforms 2j, 2j+1 of (idx, stream, pass)
__device__ __forceinline__ void philox_pair(const Philox& key, uint32_t idx, uint32_t stream,
                                            uint64_t pass, int j, double& ua, double& ub) {
  uint32_t w[4];
  philox_block(idx, (stream << 8) | (uint32_t)j, (uint32_t)pass, (uint32_t)(pass >> 32), key.k0,
               key.k1, w);
  ua = u01(w[0], w[1]);
  ub = u01(w[2], w[3]);
}

// ------------------------------------------------------------ reductions

// Deterministic in-block sum of NQ doubles (blockDim = 256 = 4 waves of 64):
// xor-butterfly inside each wave, then the four wave totals in fixed order.
// Every thread of the block receives the NQ totals.
// Deterministic block sum over NW waves (fixed butterfly and combine order);
// the result is left in every thread.
template <int NQ, int NW = 4>
__device__ __forceinline__ void block_sum(double (&v)[NQ]) {
  __shared__ double red[NW][NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double t = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    v[q] = t;
  }
  int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) red[wv][q] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double t = red[0][q];
#pragma unroll
    for (int w = 1; w < NW; ++w) t += red[w][q];
    v[q] = t;
  }
  __syncthreads();
}
template <int NQ, int NW = 4>
__device__ __forceinline__ void block_sum_store(double (&v)[NQ], double* __restrict__ out,
                                                int stride, int blk) {
  block_sum<NQ, NW>(v);
  if (threadIdx.x < NQ) {
    int q = threadIdx.x;
    double t = v[0];
#pragma unroll
    for (int k = 1; k < NQ; ++k)
      if (q == k) t = v[k];
    out[(size_t)q * stride + blk] = t;
  }
}

// Blocks are dealt round-robin to the 8 XCDs; hand each XCD a contiguous
// range of logical blocks so that neighbour gathers (z-neighbours are
// lx*ly/2 sites away) stay in that XCD's L2.
__device__ __forceinline__ int xcd_block(int nb) {
  int i = blockIdx.x;
  int k = i & 7, j = i >> 3;
  int q = nb >> 3, r = nb & 7;
  return k * q + (k < r ? k : r) + j;
}

}  // namespace lceq
