// Finite-field arithmetic for BLS12-381 on gfx950: Fp (381-bit, 14 x 28-bit
// limbs) and Fr (255-bit, 8 x 32-bit limbs) in Montgomery form, and the
// Fp2 / Fp6 / Fp12 tower used by the pairing.
//
// One lane owns one field element.  Fp multiplication is product scanning on
// 28-bit limbs with 64-bit column accumulators (see the Fp section); Fr uses
// CIOS on 32-bit limbs with the "no final carry" variant (r leaves a spare top
// bit), every inner step one v_mad_u64_u32 plus a carry add, values fully
// reduced in [0, r).
//
// Tower: Fp2 = Fp[u]/(u^2+1), Fp6 = Fp2[v]/(v^3 - xi), xi = 1+u,
//        Fp12 = Fp6[w]/(w^2 - v).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bls12_381_consts.hpp"

#define DEV __device__ __forceinline__
#define DEVI __device__ inline
#define DEVN __device__ __noinline__

struct Fp { uint32_t v[14]; };
struct Fr { uint32_t v[8]; };
struct Fp2 { Fp c0, c1; };
struct Fp6 { Fp2 c0, c1, c2; };
struct Fp12 { Fp6 c0, c1; };

// ------------------------------------------------------------------ generic limb helpers
template <int N>
DEV void mont_mul_n(uint32_t* __restrict__ r, const uint32_t* a, const uint32_t* b,
                    const uint32_t* M, uint32_t minv) {
  uint32_t t[N];
#pragma unroll
  for (int k = 0; k < N; k++) t[k] = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const uint32_t bi = b[i];
    uint64_t x = (uint64_t)a[0] * bi + t[0];
    uint32_t A = (uint32_t)(x >> 32);
    const uint32_t t0 = (uint32_t)x;
    const uint32_t m = t0 * minv;
    uint64_t y = (uint64_t)m * M[0] + t0;
    uint32_t C = (uint32_t)(y >> 32);
#pragma unroll
    for (int j = 1; j < N; j++) {
      x = (uint64_t)a[j] * bi + t[j] + A;
      A = (uint32_t)(x >> 32);
      y = (uint64_t)m * M[j] + (uint32_t)x + C;
      C = (uint32_t)(y >> 32);
      t[j - 1] = (uint32_t)y;
    }
    t[N - 1] = C + A;
  }
  // conditional subtraction of M
  uint32_t d[N];
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t s = (uint64_t)t[k] - M[k] - borrow;
    d[k] = (uint32_t)s;
    borrow = (uint32_t)(s >> 63);
  }
#pragma unroll
  for (int k = 0; k < N; k++) r[k] = borrow ? t[k] : d[k];
}

template <int N>
DEV void mod_add_n(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* M) {
  uint32_t s[N], d[N];
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t x = (uint64_t)a[k] + b[k] + c;
    s[k] = (uint32_t)x;
    c = (uint32_t)(x >> 32);
  }
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t x = (uint64_t)s[k] - M[k] - borrow;
    d[k] = (uint32_t)x;
    borrow = (uint32_t)(x >> 63);
  }
#pragma unroll
  for (int k = 0; k < N; k++) r[k] = borrow ? s[k] : d[k];
}

template <int N>
DEV void mod_sub_n(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* M) {
  uint32_t d[N];
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t x = (uint64_t)a[k] - b[k] - borrow;
    d[k] = (uint32_t)x;
    borrow = (uint32_t)(x >> 63);
  }
  const uint32_t mask = 0u - borrow;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t x = (uint64_t)d[k] + (M[k] & mask) + c;
    r[k] = (uint32_t)x;
    c = (uint32_t)(x >> 32);
  }
}

template <int N>
DEV bool limbs_is_zero(const uint32_t* a) {
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < N; k++) o |= a[k];
  return o == 0;
}

template <int N>
DEV bool limbs_eq(const uint32_t* a, const uint32_t* b) {
  uint32_t o = 0;
#pragma unroll
  for (int k = 0; k < N; k++) o |= a[k] ^ b[k];
  return o == 0;
}

// a < b (unsigned multi-precision)
template <int N>
DEV bool limbs_lt(const uint32_t* a, const uint32_t* b) {
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t x = (uint64_t)a[k] - b[k] - borrow;
    borrow = (uint32_t)(x >> 63);
  }
  return borrow != 0;
}

// ------------------------------------------------------------------ Fp
// Fp elements are 14 limbs of 28 bits (little-endian), Montgomery radix
// R = 2^392.  Invariant: every limb < 2^28 and the value < 2p.  Since 4p < R/2^9,
// a Montgomery product of two values below 2p lands below 2p without a final
// subtraction, and each column of the product (at most 28 products of < 2^56
// plus a < 2^37 carry) fits a 64-bit accumulator: one v_mad_u64_u32 per
// partial product and no carry instructions.  Values are made canonical
// (< p) only where they are compared or encoded.
static constexpr int FPW = 14;             // 32-bit words per Fp element
static constexpr uint32_t FP_MASK = (1u << 28) - 1;

DEV void fp_mul(Fp& r, const Fp& a, const Fp& b) {
  const uint32_t* M = bls::P28;
  uint32_t m[14];
  uint64_t carry = 0;
#pragma unroll
  for (int s = 0; s < 14; s++) {
    uint64_t x = carry, y = 0;
#pragma unroll
    for (int i = 0; i <= s; i++) x += (uint64_t)a.v[i] * b.v[s - i];
#pragma unroll
    for (int i = 0; i < s; i++) y += (uint64_t)m[i] * M[s - i];
    x += y;
    m[s] = ((uint32_t)x * bls::P28_INV) & FP_MASK;
    x += (uint64_t)m[s] * M[0];
    carry = x >> 28;
  }
#pragma unroll
  for (int s = 14; s < 27; s++) {
    uint64_t x = carry, y = 0;
#pragma unroll
    for (int i = s - 13; i < 14; i++) {
      x += (uint64_t)a.v[i] * b.v[s - i];
      y += (uint64_t)m[i] * M[s - i];
    }
    x += y;
    r.v[s - 14] = (uint32_t)x & FP_MASK;
    carry = x >> 28;
  }
  r.v[13] = (uint32_t)carry;
}
// Squaring: the off-diagonal products are formed once and doubled.
DEV void fp_sqr(Fp& r, const Fp& a) {
  const uint32_t* M = bls::P28;
  uint32_t m[14];
  uint64_t carry = 0;
#pragma unroll
  for (int s = 0; s < 14; s++) {
    uint64_t c = 0, y = 0;
#pragma unroll
    for (int i = 0; 2 * i < s; i++) c += (uint64_t)a.v[i] * a.v[s - i];
    uint64_t x = carry + (c << 1);
    if ((s & 1) == 0) x += (uint64_t)a.v[s / 2] * a.v[s / 2];
#pragma unroll
    for (int i = 0; i < s; i++) y += (uint64_t)m[i] * M[s - i];
    x += y;
    m[s] = ((uint32_t)x * bls::P28_INV) & FP_MASK;
    x += (uint64_t)m[s] * M[0];
    carry = x >> 28;
  }
#pragma unroll
  for (int s = 14; s < 27; s++) {
    uint64_t c = 0, y = 0;
#pragma unroll
    for (int i = s - 13; 2 * i < s; i++) c += (uint64_t)a.v[i] * a.v[s - i];
#pragma unroll
    for (int i = s - 13; i < 14; i++) y += (uint64_t)m[i] * M[s - i];
    uint64_t x = carry + (c << 1) + y;
    if ((s & 1) == 0) x += (uint64_t)a.v[s / 2] * a.v[s / 2];
    r.v[s - 14] = (uint32_t)x & FP_MASK;
    carry = x >> 28;
  }
  r.v[13] = (uint32_t)carry;
}
// r = a + b, reduced below 2p
DEV void fp_add(Fp& r, const Fp& a, const Fp& b) {
  uint32_t s[14], d[14];
  uint32_t c = 0;
  int32_t br = 0;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const uint32_t t = a.v[k] + b.v[k] + c;
    c = t >> 28;
    s[k] = t & FP_MASK;
    const int32_t u = (int32_t)s[k] - (int32_t)bls::P2_28[k] + br;
    br = u >> 28;
    d[k] = (uint32_t)u & FP_MASK;
  }
#pragma unroll
  for (int k = 0; k < 14; k++) r.v[k] = br ? s[k] : d[k];
}
// r = a - b, reduced below 2p (2p is added back when a < b)
DEV void fp_sub(Fp& r, const Fp& a, const Fp& b) {
  uint32_t d[14];
  int32_t br = 0;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const int32_t u = (int32_t)a.v[k] - (int32_t)b.v[k] + br;
    br = u >> 28;
    d[k] = (uint32_t)u & FP_MASK;
  }
  const uint32_t msk = (uint32_t)br;
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const uint32_t t = d[k] + (bls::P2_28[k] & msk) + c;
    c = t >> 28;
    r.v[k] = t & FP_MASK;
  }
}
DEV void fp_dbl(Fp& r, const Fp& a) { fp_add(r, a, a); }
// a - p if a >= p: the canonical representative of a value below 2p
DEV void fp_canon(Fp& r, const Fp& a) {
  uint32_t d[14];
  int32_t br = 0;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    const int32_t u = (int32_t)a.v[k] - (int32_t)bls::P28[k] + br;
    br = u >> 28;
    d[k] = (uint32_t)u & FP_MASK;
  }
#pragma unroll
  for (int k = 0; k < 14; k++) r.v[k] = br ? a.v[k] : d[k];
}
// zero is represented by 0 or by p
DEV bool fp_is_zero(const Fp& a) {
  uint32_t z = 0, q = 0;
#pragma unroll
  for (int k = 0; k < 14; k++) {
    z |= a.v[k];
    q |= a.v[k] ^ bls::P28[k];
  }
  return z == 0 || q == 0;
}
DEV bool fp_eq(const Fp& a, const Fp& b) {
  Fp d;
  fp_sub(d, a, b);
  return fp_is_zero(d);
}
DEV void fp_zero(Fp& r) {
#pragma unroll
  for (int k = 0; k < 14; k++) r.v[k] = 0;
}
DEV void fp_one(Fp& r) {
#pragma unroll
  for (int k = 0; k < 14; k++) r.v[k] = bls::P_ONE[k];
}
DEV void fp_neg(Fp& r, const Fp& a) {
  Fp z;
  fp_zero(z);
  fp_sub(r, z, a);
}
DEV void fp_load_const(Fp& r, const uint32_t* c) {
#pragma unroll
  for (int k = 0; k < 14; k++) r.v[k] = c[k];
}
// canonical -> Montgomery
DEV void fp_to_mont(Fp& r, const Fp& a) {
  Fp r2;
  fp_load_const(r2, bls::P_R2);
  fp_mul(r, a, r2);
}
// Montgomery -> canonical (fully reduced below p)
DEV void fp_from_mont(Fp& r, const Fp& a) {
  Fp one;
  fp_zero(one);
  one.v[0] = 1;
  fp_mul(r, a, one);
  fp_canon(r, r);
}
// r = a^e, e given as 12 little-endian 32-bit words (a constant array)
DEVN void fp_pow_const(Fp& r, const Fp& a, const uint32_t* e) {
  Fp acc;
  fp_one(acc);
  bool started = false;
#pragma unroll 1
  for (int i = 12 * 32 - 1; i >= 0; i--) {
    if (started) fp_sqr(acc, acc);
    if ((e[i >> 5] >> (i & 31)) & 1) {
      if (started)
        fp_mul(acc, acc, a);
      else
        acc = a;
      started = true;
    }
  }
  r = acc;
}
// ------------------------------------------------------------------ inversion
// Bernstein-Yang safegcd with half-delta divsteps.  f = p, g = a (canonical);
// each of 30 batches runs 30 divsteps on the low 32 bits of f and g, which
// fix the batch's transition matrix [u v; q r] (|entries| <= 2^30), then
// applies it to the full f, g (an exact division by 2^30) and to the
// coefficients d, e with f = d a, g = e a (mod p), divided by 2^30 by adding
// the multiple of p that clears the low 30 bits and kept in [0, p).  900
// divsteps (>= 879, the bound for 381-bit inputs) leave g = 0 and f = +-1, so
// a^-1 = +-d.  Numbers are 13 limbs of 30 bits, the top limb signed.  The
// divsteps are branchless and the iteration count fixed, so the lanes of a
// wave stay together.  For Montgomery inputs a = x R the result is
// multiplied by R^3 (giving x^-1 R).  a = 0 gives 0.
namespace inv30 {
constexpr int L = 13;
constexpr int32_t M30 = (1 << 30) - 1;

DEV void divsteps30(int32_t& delta2, uint32_t f, uint32_t g, int32_t& u, int32_t& v, int32_t& q,
                    int32_t& r) {
  u = 1;
  v = 0;
  q = 0;
  r = 1;
#pragma unroll
  for (int i = 0; i < 30; i++) {
    const int32_t c2 = -(int32_t)(g & 1);                // g odd
    const int32_t c1 = (delta2 > 0 ? -1 : 0) & c2;       // swap f and g
    const uint32_t x = (f ^ (uint32_t)c1) - (uint32_t)c1;  // c1 ? -f : f
    const int32_t y = (u ^ c1) - c1, z = (v ^ c1) - c1;
    g += x & (uint32_t)c2;
    q += y & c2;
    r += z & c2;
    delta2 = c1 ? 2 - delta2 : 2 + delta2;
    f += g & (uint32_t)c1;
    u += q & c1;
    v += r & c1;
    g >>= 1;
    u = (int32_t)((uint32_t)u << 1);
    v = (int32_t)((uint32_t)v << 1);
  }
}

// (x, y) <- ((u x + v y) / 2^30, (q x + r y) / 2^30), exact
DEV void update_fg(int32_t* x, int32_t* y, int32_t u, int32_t v, int32_t q, int32_t r) {
  int64_t cx = (int64_t)u * x[0] + (int64_t)v * y[0];
  int64_t cy = (int64_t)q * x[0] + (int64_t)r * y[0];
  cx >>= 30;
  cy >>= 30;
#pragma unroll
  for (int i = 1; i < L; i++) {
    cx += (int64_t)u * x[i] + (int64_t)v * y[i];
    cy += (int64_t)q * x[i] + (int64_t)r * y[i];
    x[i - 1] = (int32_t)cx & M30;
    y[i - 1] = (int32_t)cy & M30;
    cx >>= 30;
    cy >>= 30;
  }
  x[L - 1] = (int32_t)cx;
  y[L - 1] = (int32_t)cy;
}

// a += p & mask (carries propagated; top limb keeps the sign)
DEV void add_p_masked(int32_t* a, int32_t mask) {
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; i++) {
    c += a[i] + (bls::P_S30[i] & mask);
    a[i] = c & M30;
    c >>= 30;
  }
  a[L - 1] += c + (bls::P_S30[L - 1] & mask);
}

// a in (-p, 2p) -> [0, p)
DEV void normalize(int32_t* a) {
  add_p_masked(a, a[L - 1] >> 31);
  int32_t b[L];
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < L - 1; i++) {
    c += a[i] - bls::P_S30[i];
    b[i] = c & M30;
    c >>= 30;
  }
  b[L - 1] = a[L - 1] - bls::P_S30[L - 1] + c;
  const int32_t keep = b[L - 1] >> 31;  // a - p < 0: keep a
#pragma unroll
  for (int i = 0; i < L; i++) a[i] = keep ? a[i] : b[i];
}

// (d, e) <- ((u d + v e) / 2^30, (q d + r e) / 2^30) mod p, d, e in [0, p)
DEV void update_de(int32_t* d, int32_t* e, int32_t u, int32_t v, int32_t q, int32_t r) {
  int64_t cd = (int64_t)u * d[0] + (int64_t)v * e[0];
  int64_t ce = (int64_t)q * d[0] + (int64_t)r * e[0];
  const int32_t md = (int32_t)(((uint32_t)cd * bls::P_INV30) & (uint32_t)M30);
  const int32_t me = (int32_t)(((uint32_t)ce * bls::P_INV30) & (uint32_t)M30);
  cd += (int64_t)md * bls::P_S30[0];
  ce += (int64_t)me * bls::P_S30[0];
  cd >>= 30;
  ce >>= 30;
#pragma unroll
  for (int i = 1; i < L; i++) {
    cd += (int64_t)u * d[i] + (int64_t)v * e[i] + (int64_t)md * bls::P_S30[i];
    ce += (int64_t)q * d[i] + (int64_t)r * e[i] + (int64_t)me * bls::P_S30[i];
    d[i - 1] = (int32_t)cd & M30;
    e[i - 1] = (int32_t)ce & M30;
    cd >>= 30;
    ce >>= 30;
  }
  d[L - 1] = (int32_t)cd;
  e[L - 1] = (int32_t)ce;
  normalize(d);
  normalize(e);
}

// canonical 28-bit limbs -> 30-bit limbs (bit 30j starts in limb 30j/28 at
// offset 2j mod 28 <= 24, so two source limbs always cover it)
DEV void from_fp(int32_t* o, const Fp& a) {
#pragma unroll
  for (int j = 0; j < L; j++) {
    const int bit = 30 * j, i0 = bit / 28, s = bit % 28;
    uint32_t w = a.v[i0] >> s;
    if (i0 + 1 < 14) w |= a.v[i0 + 1] << (28 - s);
    o[j] = (int32_t)(w & (uint32_t)M30);
  }
}
// non-negative 30-bit limbs -> 28-bit limbs
DEV void to_fp(Fp& r, const int32_t* d) {
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const int bit = 28 * i, j0 = bit / 30, s = bit % 30;
    uint32_t w = j0 < L ? (uint32_t)d[j0] >> s : 0u;
    if (j0 + 1 < L) w |= (uint32_t)d[j0 + 1] << (30 - s);
    r.v[i] = w & FP_MASK;
  }
}
}  // namespace inv30

DEVN void fp_inv(Fp& r, const Fp& a) {
  using namespace inv30;
  Fp c;
  fp_canon(c, a);
  int32_t f[L], g[L], d[L], e[L];
#pragma unroll
  for (int i = 0; i < L; i++) {
    f[i] = bls::P_S30[i];
    d[i] = 0;
    e[i] = 0;
  }
  e[0] = 1;
  from_fp(g, c);
  int32_t delta2 = 1;
#pragma unroll 1
  for (int b = 0; b < 30; b++) {
    int32_t u, v, q, rr;
    divsteps30(delta2, (uint32_t)f[0] | ((uint32_t)f[1] << 30),
               (uint32_t)g[0] | ((uint32_t)g[1] << 30), u, v, q, rr);
    update_fg(f, g, u, v, q, rr);
    update_de(d, e, u, v, q, rr);
  }
  // f = -1: the inverse is p - d (p when d = 0, still below 2p)
  const int32_t neg = f[L - 1] >> 31;
  int32_t cy = 0;
#pragma unroll
  for (int i = 0; i < L - 1; i++) {
    cy += (bls::P_S30[i] - d[i]);
    const int32_t nd = cy & M30;
    cy >>= 30;
    d[i] = neg ? nd : d[i];
  }
  d[L - 1] = neg ? bls::P_S30[L - 1] - d[L - 1] + cy : d[L - 1];
  Fp x, r3;
  to_fp(x, d);
  fp_load_const(r3, bls::P_R3);
  fp_mul(r, x, r3);
}
// r = a^e for e given as a sliding-window schedule with 4-bit odd digits
// (first digit, then `squarings, digit` steps, then `tail` squarings); the
// odd powers a, a^3, .., a^15 are picked by selects (no indexed registers).
DEVN void fp_pow_window4(Fp& r, const Fp& a, const uint8_t* sq, const uint8_t* dig, int steps,
                         int tail) {
  Fp T[8], a2;
  fp_sqr(a2, a);
  T[0] = a;
#pragma unroll
  for (int j = 1; j < 8; j++) fp_mul(T[j], T[j - 1], a2);
  auto pick = [&](uint32_t d) {
    Fp t = T[0];
#pragma unroll
    for (int j = 1; j < 8; j++)
      if ((d >> 1) == (uint32_t)j) t = T[j];
    return t;
  };
  Fp acc = pick(dig[0]);
#pragma unroll 1
  for (int i = 1; i < steps; i++) {
#pragma unroll 1
    for (int k = 0; k < sq[i]; k++) fp_sqr(acc, acc);
    const Fp t = pick(dig[i]);
    fp_mul(acc, acc, t);
  }
#pragma unroll 1
  for (int k = 0; k < tail; k++) fp_sqr(acc, acc);
  r = acc;
}

// returns false if a is not a square
DEVI bool fp_sqrt(Fp& r, const Fp& a) {
  Fp s, t;
  fp_pow_window4(s, a, bls::SQRT_W4_SQ, bls::SQRT_W4_DIG, bls::SQRT_W4_STEPS, bls::SQRT_W4_TAIL);
  fp_sqr(t, s);
  r = s;
  return fp_eq(t, a);
}
// lexicographically largest (canonical value > (p-1)/2)
DEV bool fp_lex_largest(const Fp& a) {
  Fp c;
  fp_from_mont(c, a);
  return limbs_lt<14>(bls::P_HALF, c.v);
}
DEV uint32_t fp_sgn0(const Fp& a) {
  Fp c;
  fp_from_mont(c, a);
  return c.v[0] & 1;
}
// 384-bit little-endian words (12 x 32) <-> 28-bit limbs
DEV void fp_limbs_from_words(uint32_t* r14, const uint32_t* w12) {
#pragma unroll
  for (int i = 0; i < 14; i++) {
    const int bit = 28 * i, k = bit >> 5, sh = bit & 31;
    uint32_t v = k < 12 ? (w12[k] >> sh) : 0u;
    if (sh > 4 && k + 1 < 12) v |= w12[k + 1] << (32 - sh);
    r14[i] = v & FP_MASK;
  }
}
DEV void fp_words_from_limbs(uint32_t* w12, const uint32_t* r14) {
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const int bit = 32 * k, i = bit / 28, sh = bit % 28;
    uint32_t v = r14[i] >> sh;
    v |= r14[i + 1] << (28 - sh);
    if (sh > 24 && i + 2 < 14) v |= r14[i + 2] << (56 - sh);
    w12[k] = v;
  }
}

// ------------------------------------------------------------------ Fr
DEV void fr_mul(Fr& r, const Fr& a, const Fr& b) {
  mont_mul_n<8>(r.v, a.v, b.v, bls::R_MOD, bls::R_INV);
}
DEV void fr_add(Fr& r, const Fr& a, const Fr& b) { mod_add_n<8>(r.v, a.v, b.v, bls::R_MOD); }
DEV void fr_sub(Fr& r, const Fr& a, const Fr& b) { mod_sub_n<8>(r.v, a.v, b.v, bls::R_MOD); }
DEV bool fr_is_zero(const Fr& a) { return limbs_is_zero<8>(a.v); }
DEV void fr_zero(Fr& r) {
#pragma unroll
  for (int k = 0; k < 8; k++) r.v[k] = 0;
}
DEV void fr_neg(Fr& r, const Fr& a) {
  Fr z;
  fr_zero(z);
  fr_sub(r, z, a);
}
DEV void fr_to_mont(Fr& r, const Fr& a) {
  Fr r2;
#pragma unroll
  for (int k = 0; k < 8; k++) r2.v[k] = bls::R_R2[k];
  fr_mul(r, a, r2);
}
DEV void fr_from_mont(Fr& r, const Fr& a) {
  Fr one;
  fr_zero(one);
  one.v[0] = 1;
  fr_mul(r, a, one);
}
DEVN void fr_inv(Fr& r, const Fr& a) {  // Montgomery in, Montgomery out
  // r - 2 has its top set bit at position 254
  Fr acc = a;
#pragma unroll 1
  for (int i = 253; i >= 0; i--) {
    fr_mul(acc, acc, acc);
    if ((bls::R_MINUS_2[i >> 5] >> (i & 31)) & 1) fr_mul(acc, acc, a);
  }
  r = acc;
}
DEV bool fr_canonical(const Fr& a) { return limbs_lt<8>(a.v, bls::R_MOD); }

// Reduce a 512-bit little-endian value (16 limbs) mod r, result canonical
// (not Montgomery).  v = hi * 2^256 + lo with hi, lo < 2^256 < 3r.
DEV void fr_sub_if_geq(Fr& a) {
  uint32_t d[8];
  uint32_t borrow = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint64_t x = (uint64_t)a.v[k] - bls::R_MOD[k] - borrow;
    d[k] = (uint32_t)x;
    borrow = (uint32_t)(x >> 63);
  }
#pragma unroll
  for (int k = 0; k < 8; k++) a.v[k] = borrow ? a.v[k] : d[k];
}
DEV void fr_reduce_wide(Fr& out, const uint32_t* v16) {
  Fr hi, lo, r2, t;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    lo.v[k] = v16[k];
    hi.v[k] = v16[8 + k];
    r2.v[k] = bls::R_R2[k];
  }
  fr_sub_if_geq(hi);
  fr_sub_if_geq(hi);
  fr_sub_if_geq(lo);
  fr_sub_if_geq(lo);
  // mont(hi, R2) = hi * 2^256 mod r, canonical
  fr_mul(t, hi, r2);
  fr_add(out, t, lo);
}

// Reduce a 512-bit little-endian value (16 x 32-bit words) mod p, result in
// Montgomery form: v = hi * 2^384 + lo with lo < 2^384 (below 10p, which the
// Montgomery product accepts) and hi < 2^128.
DEVN void fp_reduce_wide_to_mont(Fp& out, const uint32_t* v16) {
  Fp lo, hi, c, t;
  fp_limbs_from_words(lo.v, v16);
  uint32_t w[12];
#pragma unroll
  for (int k = 0; k < 12; k++) w[k] = k < 4 ? v16[12 + k] : 0u;
  fp_limbs_from_words(hi.v, w);
  fp_load_const(c, bls::P_R2);
  fp_mul(t, lo, c);                 // lo * R mod p
  fp_load_const(c, bls::P_WIDE_HI);
  fp_mul(out, hi, c);               // hi * 2^384 * R mod p
  fp_add(out, out, t);
}

// ------------------------------------------------------------------ Fp2
DEV void fp2_add(Fp2& r, const Fp2& a, const Fp2& b) { fp_add(r.c0, a.c0, b.c0); fp_add(r.c1, a.c1, b.c1); }
DEV void fp2_sub(Fp2& r, const Fp2& a, const Fp2& b) { fp_sub(r.c0, a.c0, b.c0); fp_sub(r.c1, a.c1, b.c1); }
DEV void fp2_dbl(Fp2& r, const Fp2& a) { fp_dbl(r.c0, a.c0); fp_dbl(r.c1, a.c1); }
DEV void fp2_neg(Fp2& r, const Fp2& a) { fp_neg(r.c0, a.c0); fp_neg(r.c1, a.c1); }
DEV void fp2_conj(Fp2& r, const Fp2& a) { r.c0 = a.c0; fp_neg(r.c1, a.c1); }
DEV bool fp2_is_zero(const Fp2& a) { return fp_is_zero(a.c0) && fp_is_zero(a.c1); }
DEV bool fp2_eq(const Fp2& a, const Fp2& b) { return fp_eq(a.c0, b.c0) && fp_eq(a.c1, b.c1); }
DEV void fp2_zero(Fp2& r) { fp_zero(r.c0); fp_zero(r.c1); }
DEV void fp2_one(Fp2& r) { fp_one(r.c0); fp_zero(r.c1); }
DEVN void fp2_mul(Fp2& r, const Fp2& a, const Fp2& b) {
  Fp v0, v1, s0, s1;
  fp_mul(v0, a.c0, b.c0);
  fp_mul(v1, a.c1, b.c1);
  fp_add(s0, a.c0, a.c1);
  fp_add(s1, b.c0, b.c1);
  fp_mul(s0, s0, s1);
  fp_sub(r.c0, v0, v1);
  fp_sub(s0, s0, v0);
  fp_sub(r.c1, s0, v1);
}
DEVN void fp2_sqr(Fp2& r, const Fp2& a) {
  Fp s, d, m;
  fp_add(s, a.c0, a.c1);
  fp_sub(d, a.c0, a.c1);
  fp_mul(m, a.c0, a.c1);
  fp_mul(r.c0, s, d);
  fp_dbl(r.c1, m);
}
DEV void fp2_mul_fp(Fp2& r, const Fp2& a, const Fp& b) { fp_mul(r.c0, a.c0, b); fp_mul(r.c1, a.c1, b); }
// multiply by xi = 1 + u
DEV void fp2_mul_xi(Fp2& r, const Fp2& a) {
  Fp t0, t1;
  fp_sub(t0, a.c0, a.c1);
  fp_add(t1, a.c0, a.c1);
  r.c0 = t0;
  r.c1 = t1;
}
DEVN void fp2_inv(Fp2& r, const Fp2& a) {
  Fp n, t;
  fp_sqr(n, a.c0);
  fp_sqr(t, a.c1);
  fp_add(n, n, t);
  fp_inv(n, n);
  fp_mul(r.c0, a.c0, n);
  fp_mul(t, a.c1, n);
  fp_neg(r.c1, t);
}
DEVN void fp2_pow_const(Fp2& r, const Fp2& a, const uint32_t* e) {
  Fp2 acc;
  fp2_one(acc);
#pragma unroll 1
  for (int i = 12 * 32 - 1; i >= 0; i--) {
    fp2_sqr(acc, acc);
    if ((e[i >> 5] >> (i & 31)) & 1) fp2_mul(acc, acc, a);
  }
  r = acc;
}
DEV void fp2_load_const(Fp2& r, const uint32_t* c) {
  fp_load_const(r.c0, c);
  fp_load_const(r.c1, c + FPW);
}
// Square root in Fp2, p = 3 mod 4 (Adj & Rodriguez-Henriquez, Algorithm 9).
DEVN bool fp2_sqrt(Fp2& r, const Fp2& a) {
  if (fp2_is_zero(a)) {
    r = a;
    return true;
  }
  Fp2 a1, alpha, x0, t, one, minus_one;
  fp2_one(one);
  fp2_neg(minus_one, one);
  fp2_pow_const(a1, a, bls::P_MINUS3_DIV4);
  fp2_mul(x0, a1, a);
  fp2_mul(alpha, a1, x0);
  if (fp2_eq(alpha, minus_one)) {
    // x = u * x0
    fp_neg(t.c0, x0.c1);
    t.c1 = x0.c0;
  } else {
    Fp2 b;
    fp2_add(b, alpha, one);
    fp2_pow_const(b, b, bls::P_LEGENDRE_EXP);
    fp2_mul(t, b, x0);
  }
  Fp2 chk;
  fp2_sqr(chk, t);
  r = t;
  return fp2_eq(chk, a);
}
DEV bool fp2_lex_largest(const Fp2& a) {
  if (!fp_is_zero(a.c1)) return fp_lex_largest(a.c1);
  return fp_lex_largest(a.c0);
}

// ------------------------------------------------------------------ Fp6
DEV void fp6_add(Fp6& r, const Fp6& a, const Fp6& b) { fp2_add(r.c0, a.c0, b.c0); fp2_add(r.c1, a.c1, b.c1); fp2_add(r.c2, a.c2, b.c2); }
DEV void fp6_sub(Fp6& r, const Fp6& a, const Fp6& b) { fp2_sub(r.c0, a.c0, b.c0); fp2_sub(r.c1, a.c1, b.c1); fp2_sub(r.c2, a.c2, b.c2); }
DEV void fp6_neg(Fp6& r, const Fp6& a) { fp2_neg(r.c0, a.c0); fp2_neg(r.c1, a.c1); fp2_neg(r.c2, a.c2); }
DEV void fp6_zero(Fp6& r) { fp2_zero(r.c0); fp2_zero(r.c1); fp2_zero(r.c2); }
DEVN void fp6_mul(Fp6& r, const Fp6& a, const Fp6& b) {
  Fp2 t0, t1, t2, s0, s1, c0, c1, c2;
  fp2_mul(t0, a.c0, b.c0);
  fp2_mul(t1, a.c1, b.c1);
  fp2_mul(t2, a.c2, b.c2);
  fp2_add(s0, a.c1, a.c2);
  fp2_add(s1, b.c1, b.c2);
  fp2_mul(c0, s0, s1);
  fp2_sub(c0, c0, t1);
  fp2_sub(c0, c0, t2);
  fp2_mul_xi(c0, c0);
  fp2_add(c0, c0, t0);
  fp2_add(s0, a.c0, a.c1);
  fp2_add(s1, b.c0, b.c1);
  fp2_mul(c1, s0, s1);
  fp2_sub(c1, c1, t0);
  fp2_sub(c1, c1, t1);
  fp2_mul_xi(s0, t2);
  fp2_add(c1, c1, s0);
  fp2_add(s0, a.c0, a.c2);
  fp2_add(s1, b.c0, b.c2);
  fp2_mul(c2, s0, s1);
  fp2_sub(c2, c2, t0);
  fp2_sub(c2, c2, t2);
  fp2_add(r.c2, c2, t1);
  r.c0 = c0;
  r.c1 = c1;
}
// Chung-Hasan SQR2
DEVN void fp6_sqr(Fp6& r, const Fp6& a) {
  Fp2 s0, s1, s2, s3, s4, t;
  fp2_sqr(s0, a.c0);
  fp2_mul(s1, a.c0, a.c1);
  fp2_dbl(s1, s1);
  fp2_sub(t, a.c0, a.c1);
  fp2_add(t, t, a.c2);
  fp2_sqr(s2, t);
  fp2_mul(s3, a.c1, a.c2);
  fp2_dbl(s3, s3);
  fp2_sqr(s4, a.c2);
  // c0 = s0 + xi s3 ; c1 = s1 + xi s4 ; c2 = s1 + s2 + s3 - s0 - s4
  Fp2 c0, c1, c2;
  fp2_mul_xi(t, s3);
  fp2_add(c0, s0, t);
  fp2_mul_xi(t, s4);
  fp2_add(c1, s1, t);
  fp2_add(c2, s1, s2);
  fp2_add(c2, c2, s3);
  fp2_sub(c2, c2, s0);
  fp2_sub(c2, c2, s4);
  r.c0 = c0;
  r.c1 = c1;
  r.c2 = c2;
}
DEV void fp6_mul_v(Fp6& r, const Fp6& a) {
  Fp2 t;
  fp2_mul_xi(t, a.c2);
  r.c2 = a.c1;
  r.c1 = a.c0;
  r.c0 = t;
}
DEVN void fp6_inv(Fp6& r, const Fp6& a) {
  Fp2 t0, t1, t2, tmp, d;
  fp2_sqr(t0, a.c0);
  fp2_mul(tmp, a.c1, a.c2);
  fp2_mul_xi(tmp, tmp);
  fp2_sub(t0, t0, tmp);
  fp2_sqr(t1, a.c2);
  fp2_mul_xi(t1, t1);
  fp2_mul(tmp, a.c0, a.c1);
  fp2_sub(t1, t1, tmp);
  fp2_sqr(t2, a.c1);
  fp2_mul(tmp, a.c0, a.c2);
  fp2_sub(t2, t2, tmp);
  fp2_mul(d, a.c2, t1);
  fp2_mul(tmp, a.c1, t2);
  fp2_add(d, d, tmp);
  fp2_mul_xi(d, d);
  fp2_mul(tmp, a.c0, t0);
  fp2_add(d, d, tmp);
  fp2_inv(d, d);
  fp2_mul(r.c0, t0, d);
  fp2_mul(r.c1, t1, d);
  fp2_mul(r.c2, t2, d);
}

// ------------------------------------------------------------------ Fp12
DEV void fp12_one(Fp12& r) {
  fp6_zero(r.c0);
  fp6_zero(r.c1);
  fp_one(r.c0.c0.c0);
}
DEVN bool fp12_is_one(const Fp12& a) {
  Fp one;
  fp_one(one);
  bool ok = fp_eq(a.c0.c0.c0, one) && fp_is_zero(a.c0.c0.c1);
  ok = ok && fp2_is_zero(a.c0.c1) && fp2_is_zero(a.c0.c2);
  ok = ok && fp2_is_zero(a.c1.c0) && fp2_is_zero(a.c1.c1) && fp2_is_zero(a.c1.c2);
  return ok;
}
DEVN void fp12_mul(Fp12& r, const Fp12& a, const Fp12& b) {
  Fp6 t0, t1, s0, s1;
  fp6_mul(t0, a.c0, b.c0);
  fp6_mul(t1, a.c1, b.c1);
  fp6_add(s0, a.c0, a.c1);
  fp6_add(s1, b.c0, b.c1);
  fp6_mul(s0, s0, s1);
  fp6_sub(s0, s0, t0);
  fp6_sub(r.c1, s0, t1);
  fp6_mul_v(t1, t1);
  fp6_add(r.c0, t0, t1);
}
// complex squaring: (a0 + a1 w)^2 = (a0^2 + v a1^2) + 2 a0 a1 w
DEVN void fp12_sqr(Fp12& r, const Fp12& a) {
  Fp6 t0, t1, s;
  fp6_mul(t0, a.c0, a.c1);
  fp6_add(s, a.c0, a.c1);
  fp6_mul_v(t1, a.c1);
  fp6_add(t1, a.c0, t1);
  fp6_mul(s, s, t1);          // (a0 + a1)(a0 + v a1) = a0^2 + v a1^2 + (1+v) a0 a1
  fp6_sub(s, s, t0);
  fp6_mul_v(t1, t0);
  fp6_sub(r.c0, s, t1);
  fp6_add(r.c1, t0, t0);
}
DEV void fp12_conj(Fp12& r, const Fp12& a) {
  r.c0 = a.c0;
  fp6_neg(r.c1, a.c1);
}
DEVN void fp12_inv(Fp12& r, const Fp12& a) {
  Fp6 t0, t1;
  fp6_sqr(t0, a.c0);
  fp6_sqr(t1, a.c1);
  fp6_mul_v(t1, t1);
  fp6_sub(t0, t0, t1);
  fp6_inv(t0, t0);
  fp6_mul(r.c0, a.c0, t0);
  fp6_mul(t1, a.c1, t0);
  fp6_neg(r.c1, t1);
}
// Frobenius x -> x^p.  Coefficient of w^k (k = 2i + j for c_j.c_i) is
// conjugated and multiplied by gamma1[k] = xi^(k (p-1)/6).
DEVN void fp12_frob(Fp12& r, const Fp12& a) {
  Fp2 g;
  fp2_conj(r.c0.c0, a.c0.c0);
  fp2_conj(r.c0.c1, a.c0.c1);
  fp2_load_const(g, bls::FROB1 + 2 * FPW * 2);
  fp2_mul(r.c0.c1, r.c0.c1, g);
  fp2_conj(r.c0.c2, a.c0.c2);
  fp2_load_const(g, bls::FROB1 + 2 * FPW * 4);
  fp2_mul(r.c0.c2, r.c0.c2, g);
  fp2_conj(r.c1.c0, a.c1.c0);
  fp2_load_const(g, bls::FROB1 + 2 * FPW * 1);
  fp2_mul(r.c1.c0, r.c1.c0, g);
  fp2_conj(r.c1.c1, a.c1.c1);
  fp2_load_const(g, bls::FROB1 + 2 * FPW * 3);
  fp2_mul(r.c1.c1, r.c1.c1, g);
  fp2_conj(r.c1.c2, a.c1.c2);
  fp2_load_const(g, bls::FROB1 + 2 * FPW * 5);
  fp2_mul(r.c1.c2, r.c1.c2, g);
}
// x -> x^(p^2): coefficients stay, scaled by gamma2[k] in Fp.
DEVN void fp12_frob2(Fp12& r, const Fp12& a) {
  Fp g;
  r.c0.c0 = a.c0.c0;
  fp_load_const(g, bls::FROB2 + FPW * 2);
  fp2_mul_fp(r.c0.c1, a.c0.c1, g);
  fp_load_const(g, bls::FROB2 + FPW * 4);
  fp2_mul_fp(r.c0.c2, a.c0.c2, g);
  fp_load_const(g, bls::FROB2 + FPW * 1);
  fp2_mul_fp(r.c1.c0, a.c1.c0, g);
  fp_load_const(g, bls::FROB2 + FPW * 3);
  fp2_mul_fp(r.c1.c1, a.c1.c1, g);
  fp_load_const(g, bls::FROB2 + FPW * 5);
  fp2_mul_fp(r.c1.c2, a.c1.c2, g);
}

// ------------------------------------------------------------------ byte codecs
// 48-byte big-endian -> canonical limbs; false when >= p (flags must be cleared
// by the caller).
DEV bool fp_from_be_bytes(Fp& r, const uint8_t* be) {
  uint32_t w[12];
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const uint8_t* q = be + 44 - 4 * k;
    w[k] = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3];
  }
  fp_limbs_from_words(r.v, w);
  return limbs_lt<14>(r.v, bls::P28);
}
// canonical (below p) -> 48-byte big-endian
DEV void fp_to_be_bytes(uint8_t* be, const Fp& canonical) {
  uint32_t w[12];
  fp_words_from_limbs(w, canonical.v);
#pragma unroll
  for (int k = 0; k < 12; k++) {
    uint8_t* q = be + 44 - 4 * k;
    q[0] = (uint8_t)(w[k] >> 24);
    q[1] = (uint8_t)(w[k] >> 16);
    q[2] = (uint8_t)(w[k] >> 8);
    q[3] = (uint8_t)w[k];
  }
}
DEV void fr_from_be_bytes_raw(Fr& r, const uint8_t* be) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint8_t* q = be + 28 - 4 * k;
    r.v[k] = ((uint32_t)q[0] << 24) | ((uint32_t)q[1] << 16) | ((uint32_t)q[2] << 8) | q[3];
  }
}
DEV void fr_to_be_bytes(uint8_t* be, const Fr& canonical) {
#pragma unroll
  for (int k = 0; k < 8; k++) {
    uint8_t* q = be + 28 - 4 * k;
    uint32_t w = canonical.v[k];
    q[0] = (uint8_t)(w >> 24);
    q[1] = (uint8_t)(w >> 16);
    q[2] = (uint8_t)(w >> 8);
    q[3] = (uint8_t)w;
  }
}
