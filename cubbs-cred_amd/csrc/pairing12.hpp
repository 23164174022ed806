// Pairing check on 12 lanes: one lane per Fp coefficient of Fp12.
//
// An Fp12 element is held in the basis 1, w, .., w^5 (w^6 = xi = 1 + u); lane
// (k, c) of a 12-lane group owns component c (u^0 or u^1) of the Fp2
// coefficient of w^k.  A wavefront carries 5 checks (lanes 0..59); the last
// 4 lanes run the same instruction stream on the fifth check's LDS without
// storing, and their results are discarded.
//
// Every output coefficient is a short sum of Fp products.  The products are
// accumulated unreduced in 27 64-bit columns (28-bit limbs, at most 12
// products of factors below 4p) and reduced once (Montgomery, R = 2^392):
// a full Fp12 multiplication costs 12 products + 1 reduction per lane, a
// squaring 8 + 1, a sparse line 5 + 1, a cyclotomic squaring 3 + 1.
// Negative terms and the xi factor are folded into the factors: before a
// product the operand is published to LDS together with derived values
// (2y, y0 + 2p - y1, y0 + y1, 4p - y0 - y1, 2p - y1), so a lane forms
//   plain  c=0: x0 y0 + x1 (2p - y1)      c=1: x0 y1 + x1 y0
//   xi     c=0: x0 (y0 - y1) - x1 (y0 + y1)   c=1: x0 (y0 + y1) + x1 (y0 - y1)
// with only LDS address selection, no branches.
//
// All functions contain __syncthreads(): every thread of the 64-thread block
// calls them in the same order.  Slot k maps to the tower coordinate
// c_{k mod 2}.c_{k div 2} of pairing.hpp.
#pragma once
#include "pairing.hpp"

namespace p12 {

constexpr int E = FPW;                                   // words of an Fp value
enum { EY0, EY1, E2Y0, E2Y1, ED, ES, ENS, EN1, NE };     // derived entries per slot
// LDS layout, padded so that the 60 working lanes' reads of one product term
// spread over the 32 banks: entries LE words apart, slots SLOT words apart,
// checks LDS_WORDS apart.  With the unpadded strides (14, 112, 840) the
// lanes of a check hit only a few banks (slot stride 112 = 16 mod 32), which
// cost about 3.5 LDS cycles per access against 1.37 with these (a model of
// every access pattern below over 32 banks per 32 lanes).
constexpr int LE = 15;
constexpr int SLOT = 127;
constexpr int REGION = 6 * SLOT;                         // rich entries of one operand
constexpr int RAW = 6 * 2 * LE;                          // plain coefficients of one operand
constexpr int LDS_WORDS = 949;                           // per check
static_assert(LE >= E && SLOT >= NE * LE && LDS_WORDS >= REGION + RAW, "LDS layout");
constexpr int CHECKS_PER_BLOCK = 5;
// LDS of a 64-thread block: 5 checks and one shared zero entry (the 4 spare
// lanes alias the last check's region and never store); 19,040 bytes, so the
// eight blocks of a CU at two per SIMD fit its 160 KiB
constexpr int BLOCK_LDS_WORDS = CHECKS_PER_BLOCK * LDS_WORDS + LE;
// Per check, line and pairing: the second factors of the sparse line
// product, so that the Miller loop forms no derived values.  With
// m = -lambda x_P (Fp2) and c the line's constant (Fp2):
//   M0 = m0, M1 = m1, NM1 = 2p - m1, D = m0 + 2p - m1, NS = 4p - m0 - m1,
//   S = m0 + m1, C0 = c0, C1 = c1, NC1 = 2p - c1.
// A skipped pairing (identity point or key) stores the line 1.
enum { LV_M0, LV_M1, LV_NM1, LV_D, LV_NS, LV_S, LV_C0, LV_C1, LV_NC1, LV_N };
constexpr int L2_WORDS = MILLER_LINES * 2 * LV_N * E;

// LDS words (address space 3, so accesses compile to ds_* instructions even
// across the non-inlined operations below)
typedef __attribute__((address_space(3))) uint32_t lds_t;

struct Lane {
  lds_t* F;  // rich entries of f (of g in a multiplication)
  lds_t* X;  // plain coefficients of f in a multiplication; flags of is_one
  lds_t* Z;  // a zero entry
  int k, c;
  bool wr;   // false for the spare lanes: they compute but never store
};

struct Wide {
  uint64_t c[27];
};

DEV void w_zero(Wide& w) {
#pragma unroll
  for (int i = 0; i < 27; i++) w.c[i] = 0;
}
DEV void w_mac(Wide& w, const Fp& x, const Fp& y) {
#pragma unroll
  for (int i = 0; i < 14; i++)
#pragma unroll
    for (int j = 0; j < 14; j++) w.c[i + j] += (uint64_t)x.v[i] * y.v[j];
}
// r = T / R mod p for T < p R, r < 2p.  Column sums stay below 2^64: at most
// 13 x 14 partial products of < 2^56 each per column.
DEV void w_reduce(Fp& r, Wide& w) {
  const uint32_t* M = bls::P28;
  uint64_t carry = 0;
#pragma unroll
  for (int s = 0; s < 14; s++) {
    uint64_t x = w.c[s] + carry;
    const uint32_t m = ((uint32_t)x * bls::P28_INV) & FP_MASK;
    x += (uint64_t)m * M[0];
    carry = x >> 28;
#pragma unroll
    for (int j = 1; j < 14; j++) w.c[s + j] += (uint64_t)m * M[j];
  }
#pragma unroll
  for (int s = 14; s < 27; s++) {
    const uint64_t x = w.c[s] + carry;
    r.v[s - 14] = (uint32_t)x & FP_MASK;
    carry = x >> 28;
  }
  r.v[13] = (uint32_t)carry;
}

DEV void ld(Fp& r, const uint32_t* p) {
#pragma unroll
  for (int i = 0; i < E; i++) r.v[i] = p[i];
}
DEV void st(uint32_t* p, const Fp& v) {
#pragma unroll
  for (int i = 0; i < E; i++) p[i] = v.v[i];
}
DEV void ld(Fp& r, const lds_t* p) {
#pragma unroll
  for (int i = 0; i < E; i++) r.v[i] = p[i];
}
DEV void st(lds_t* p, const Fp& v) {
#pragma unroll
  for (int i = 0; i < E; i++) p[i] = v.v[i];
}
DEV void mac_at(Wide& w, const lds_t* xa, const lds_t* ya) {
  Fp x, y;
  ld(x, xa);
  ld(y, ya);
  w_mac(w, x, y);
}
DEV void sel(Fp& r, bool t, const Fp& a, const Fp& b) {
#pragma unroll
  for (int i = 0; i < E; i++) r.v[i] = t ? a.v[i] : b.v[i];
}
// value of the adjacent lane (lane ^ 1): the other component of the slot
DEV void partner(Fp& o, const Fp& v) {
#pragma unroll
  for (int i = 0; i < E; i++) o.v[i] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v.v[i], 0xB1, 0xF, 0xF, false);
}
// r = a + b + K - s - t on normalized limbs without reduction; the caller
// guarantees a non-negative result below 2^392.  K is 0, 2p or 4p (kk).
DEV void lz5(Fp& r, const Fp& a, const Fp& b, int kk, const Fp& s, const Fp& t) {
  int32_t cy = 0;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const uint32_t K = kk == 2 ? bls::P2_28[i] : (kk == 4 ? bls::P4_28[i] : 0u);
    const int32_t u = (int32_t)(a.v[i] + b.v[i] + K) - (int32_t)s.v[i] - (int32_t)t.v[i] + cy;
    cy = u >> 28;
    r.v[i] = (uint32_t)u & FP_MASK;
  }
}

// Entry of the second factor for component c of x * y (or xi x y).
DEV int ysel0(int c, bool xi) { return c ? (xi ? ES : EY1) : (xi ? ED : EY0); }
DEV int ysel1(int c, bool xi) { return c ? (xi ? ED : EY0) : (xi ? ENS : EN1); }

// Publish the lane's coefficient v and its share of the derived entries.
// Lane c = 0 writes y0, 2 y0, y0 + 2p - y1, y0 + y1; lane c = 1 writes y1,
// 2 y1, 2p - y1, 4p - y0 - y1.  Values below 2p in, derived values below 4p.
// doubles = false leaves out 2 y0 and 2 y1, which only squarings read.
DEV void put_rich(lds_t* R, const Fp& v, const Lane& L, bool doubles = true) {
  Fp o, y0, y1, z, a, b, s2, t2;
  partner(o, v);
  sel(y0, L.c, o, v);
  sel(y1, L.c, v, o);
  fp_zero(z);
  lds_t* S = R + L.k * SLOT;
  Fp t1, t2v, t3;
  if (doubles) lz5(t1, v, v, 0, z, z);
  sel(a, L.c, z, y0);
  lz5(t2v, a, z, 2, y1, z);  // c=0: y0 + 2p - y1   c=1: 2p - y1
  sel(b, L.c, z, y1);
  sel(s2, L.c, y0, z);
  sel(t2, L.c, y1, z);
  lz5(t3, a, b, L.c ? 4 : 0, s2, t2);  // c=0: y0 + y1   c=1: 4p - y0 - y1
  if (L.wr) {
    st(S + (L.c ? EY1 : EY0) * LE, v);
    if (doubles) st(S + (L.c ? E2Y1 : E2Y0) * LE, t1);
    st(S + (L.c ? EN1 : ED) * LE, t2v);
    st(S + (L.c ? ENS : ES) * LE, t3);
  }
}

// Publish what a sparse line product reads (mul_line): y0, y1 and, one per
// lane, y0 + 2p - y1 (ED, lane c = 0) and y0 + y1 (ES, lane c = 1).  The
// other derived entries keep stale values that the line product never reads.
DEV void put_line(lds_t* R, const Fp& v, const Lane& L) {
  Fp o, y0, y1, z, b, sb, t;
  partner(o, v);
  sel(y0, L.c, o, v);
  sel(y1, L.c, v, o);
  fp_zero(z);
  sel(b, L.c, y1, z);
  sel(sb, L.c, z, y1);
  lz5(t, y0, b, L.c ? 0 : 2, sb, z);  // c=0: y0 + 2p - y1   c=1: y0 + y1
  lds_t* S = R + L.k * SLOT;
  if (L.wr) {
    st(S + (L.c ? EY1 : EY0) * LE, v);
    st(S + (L.c ? ES : ED) * LE, t);
  }
}

// f <- f * g (expanded in the loops that use it)
DEV void mul_in(Fp& f, const Fp& g, const Lane& L) {
  __syncthreads();
  put_rich(L.F, g, L, false);
  if (L.wr) st(L.X + (2 * L.k + L.c) * LE, f);
  __syncthreads();
  Wide w;
  w_zero(w);
#pragma unroll 1
  for (int i = 0; i < 6; i++) {
    int j = L.k - i;
    const bool xi = j < 0;
    j += xi ? 6 : 0;
    const lds_t* X = L.X + 2 * i * LE;
    const lds_t* Y = L.F + j * SLOT;
    mac_at(w, X, Y + ysel0(L.c, xi) * LE);
    mac_at(w, X + LE, Y + ysel1(L.c, xi) * LE);
  }
  w_reduce(f, w);
}

// f <- f^2: terms {i, j, times2, times_xi} of (f^2)_k, i < 0 = empty
__constant__ static const int8_t SQ_TERMS[6][4][4] = {
    {{0, 0, 0, 0}, {1, 5, 1, 1}, {2, 4, 1, 1}, {3, 3, 0, 1}},
    {{0, 1, 1, 0}, {2, 5, 1, 1}, {3, 4, 1, 1}, {-1, 0, 0, 0}},
    {{0, 2, 1, 0}, {1, 1, 0, 0}, {3, 5, 1, 1}, {4, 4, 0, 1}},
    {{0, 3, 1, 0}, {1, 2, 1, 0}, {4, 5, 1, 1}, {-1, 0, 0, 0}},
    {{0, 4, 1, 0}, {1, 3, 1, 0}, {2, 2, 0, 0}, {5, 5, 0, 1}},
    {{0, 5, 1, 0}, {1, 4, 1, 0}, {2, 3, 1, 0}, {-1, 0, 0, 0}}};

DEV void sqr(Fp& f, const Lane& L) {
  __syncthreads();
  put_rich(L.F, f, L);
  __syncthreads();
  Wide w;
  w_zero(w);
#pragma unroll 1
  for (int t = 0; t < 4; t++) {
    const int i = SQ_TERMS[L.k][t][0];
    const int j = SQ_TERMS[L.k][t][1];
    const bool two = SQ_TERMS[L.k][t][2] != 0;
    const bool xi = SQ_TERMS[L.k][t][3] != 0;
    const bool none = i < 0;
    const lds_t* x0 = none ? L.Z : L.F + i * SLOT + (two ? E2Y0 : EY0) * LE;
    const lds_t* x1 = none ? L.Z : L.F + i * SLOT + (two ? E2Y1 : EY1) * LE;
    const lds_t* y0 = none ? L.Z : L.F + j * SLOT + ysel0(L.c, xi) * LE;
    const lds_t* y1 = none ? L.Z : L.F + j * SLOT + ysel1(L.c, xi) * LE;
    mac_at(w, x0, y0);
    mac_at(w, x1, y1);
  }
  w_reduce(f, w);
}

// Fill one block of LV_N values (see L2_WORDS) for line constant (c0, c1)
// and m = (m0, m1), all below 2p + 1.
DEV void line_values(uint32_t* v, const Fp& m0, const Fp& m1, const Fp& c0, const Fp& c1,
                     bool skip) {
  Fp z, one, a0, a1, b0, b1, t;
  fp_zero(z);
  fp_one(one);
  sel(a0, skip, z, m0);
  sel(a1, skip, z, m1);
  sel(b0, skip, one, c0);
  sel(b1, skip, z, c1);
  st(v + LV_M0 * E, a0);
  st(v + LV_M1 * E, a1);
  lz5(t, z, z, 2, a1, z);
  st(v + LV_NM1 * E, t);
  lz5(t, a0, z, 2, a1, z);
  st(v + LV_D * E, t);
  lz5(t, z, z, 4, a0, a1);
  st(v + LV_NS * E, t);
  lz5(t, a0, a1, 0, z, z);
  st(v + LV_S * E, t);
  st(v + LV_C0 * E, b0);
  st(v + LV_C1 * E, b1);
  lz5(t, z, z, 2, b1, z);
  st(v + LV_NC1 * E, t);
}

// f <- f * (l0 + l2 w^2 + yp w^3) for one prepared line, yp in Fp (the
// evaluated Miller-loop line, see pairing.hpp), the other second factors
// read from the line's block v:
//   w^0: f_k * l0             c=0: (c0, 2p - c1)            c=1: (c1, c0)
//   w^2: f_(k-2) * l2         c=0: (m0, 2p - m1)            c=1: (m1, m0)
//        times xi when k < 2  c=0: (m0 + 2p - m1, 4p - m0 - m1)  c=1: (m0 + m1, m0 + 2p - m1)
//   w^3: f_(k-3) * yp, times xi when k < 3
// The factors are loaded before the barrier so the loads overlap put_rich.
DEV void mul_line(Fp& f, const uint32_t* v, const Fp& yp, const Lane& L) {
  const bool xi = L.k < 2;
  Fp a0, a1, b0, b1;
  ld(a0, v + (L.c ? LV_C1 : LV_C0) * E);
  ld(a1, v + (L.c ? LV_C0 : LV_NC1) * E);
  ld(b0, v + (xi ? (L.c ? LV_S : LV_D) : (L.c ? LV_M1 : LV_M0)) * E);
  ld(b1, v + (xi ? (L.c ? LV_D : LV_NS) : (L.c ? LV_M0 : LV_NM1)) * E);
  __syncthreads();
  put_line(L.F, f, L);
  __syncthreads();
  Wide w;
  w_zero(w);
  {
    Fp x0, x1;
    ld(x0, L.F + L.k * SLOT + EY0 * LE);
    ld(x1, L.F + L.k * SLOT + EY1 * LE);
    w_mac(w, x0, a0);
    w_mac(w, x1, a1);
  }
  {
    const int i = xi ? L.k + 4 : L.k - 2;
    Fp x0, x1;
    ld(x0, L.F + i * SLOT + EY0 * LE);
    ld(x1, L.F + i * SLOT + EY1 * LE);
    w_mac(w, x0, b0);
    w_mac(w, x1, b1);
  }
  {
    const bool xi3 = L.k < 3;
    const int i = xi3 ? L.k + 3 : L.k - 3;
    Fp x;
    ld(x, L.F + i * SLOT + ysel0(L.c, xi3) * LE);
    w_mac(w, x, yp);
  }
  w_reduce(f, w);
}

// Granger-Scott squaring in the cyclotomic subgroup.  Fp4 pairs (a, b) =
// (w^P, w^(P+3)); with s0 = a^2 + xi b^2 and s1 = 2ab the new coefficients are
//   w0 = 3 s0(P=0) - 2 w0   w3 = 3 s1(P=0) + 2 w3
//   w2 = 3 s0(P=1) - 2 w2   w5 = 3 s1(P=1) + 2 w5
//   w4 = 3 s0(P=2) - 2 w4   w1 = 3 xi s1(P=2) + 2 w1
// Each component of s0, s1, xi s1 is a sum of three products of derived
// entries ({which, entry, which, entry}, which: 0 = a, 1 = b, 2 = zero).
__constant__ static const int8_t GS_TERMS[3][2][3][4] = {
    // s0.c0 = (a0+a1)(a0-a1) + (b0+b1)(b0-b1) - 2 b0 b1
    // s0.c1 = 2 a0 a1 + (b0+b1)(b0-b1) + 2 b0 b1
    {{{0, ES, 0, ED}, {1, ES, 1, ED}, {1, E2Y0, 1, EN1}},
     {{0, E2Y0, 0, EY1}, {1, ES, 1, ED}, {1, E2Y0, 1, EY1}}},
    // s1 = 2ab
    {{{0, E2Y0, 1, EY0}, {0, E2Y1, 1, EN1}, {2, 0, 2, 0}},
     {{0, E2Y0, 1, EY1}, {0, E2Y1, 1, EY0}, {2, 0, 2, 0}}},
    // xi s1
    {{{0, E2Y0, 1, ED}, {0, E2Y1, 1, ENS}, {2, 0, 2, 0}},
     {{0, E2Y0, 1, ES}, {0, E2Y1, 1, ED}, {2, 0, 2, 0}}}};
__constant__ static const int8_t GS_SLOT[6][2] = {  // {P, type}
    {0, 0}, {2, 2}, {1, 0}, {0, 1}, {2, 0}, {1, 1}};

DEV void cyc_sqr(Fp& f, const Lane& L) {
  __syncthreads();
  put_rich(L.F, f, L);
  __syncthreads();
  const int P = GS_SLOT[L.k][0], ty = GS_SLOT[L.k][1];
  Wide w;
  w_zero(w);
#pragma unroll 1
  for (int t = 0; t < 3; t++) {
    const int8_t* T = GS_TERMS[ty][L.c][t];
    const lds_t* x = T[0] == 2 ? L.Z : L.F + (P + 3 * T[0]) * SLOT + T[1] * LE;
    const lds_t* y = T[2] == 2 ? L.Z : L.F + (P + 3 * T[2]) * SLOT + T[3] * LE;
    mac_at(w, x, y);
  }
  Fp s;
  w_reduce(s, w);
  // 3s - 2f (even k) or 3s + 2f (odd k) in one signed limb pass, as
  // 3s + 4p - 2f or 3s + 2f (below 10p; s, f below 2p), then minus q 2p for
  // q = floor(value / 2p) estimated from the top limb (never above the true
  // quotient, at most one below it) and one conditional subtraction of 2p:
  // the result is below 2p.  About 260 instructions instead of five modular
  // additions (~590).
  const bool even = (L.k & 1) == 0;
  Fp o;
  int32_t cy = 0;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int32_t f2 = (int32_t)(f.v[i] << 1);
    const int32_t u = (int32_t)(3u * s.v[i]) + (even ? (int32_t)bls::P4_28[i] - f2 : f2) + cy;
    cy = u >> 28;
    o.v[i] = (uint32_t)u & FP_MASK;
  }
  constexpr uint32_t T = bls::P2_28[E - 1] + 1;
  const uint32_t top = o.v[E - 1];
  const uint32_t q = (top >= T) + (top >= 2 * T) + (top >= 3 * T) + (top >= 4 * T);
  Fp d, e;
  int32_t b1 = 0, b2 = 0;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int32_t u = (int32_t)o.v[i] - (int32_t)(q * bls::P2_28[i]) + b1;
    b1 = u >> 28;
    d.v[i] = (uint32_t)u & FP_MASK;
    const int32_t t = (int32_t)d.v[i] - (int32_t)bls::P2_28[i] + b2;
    b2 = t >> 28;
    e.v[i] = (uint32_t)t & FP_MASK;
  }
  sel(f, b2 < 0, d, e);
}


// f <- conj(f) (x -> x^(p^6)): odd slots negated
DEV void conj(Fp& f, const Lane& L) {
  Fp n;
  fp_neg(n, f);
  sel(f, (L.k & 1) != 0, n, f);
}
// x -> x^p: slot k conjugated, times gamma1[k] = xi^(k (p-1)/6)
DEV void frob(Fp& f, const Lane& L) {
  const Fp v = f;
  Fp o, x0, x1, g0, g1, z, n0, y0, y1;
  partner(o, v);
  sel(x0, L.c, o, v);
  sel(x1, L.c, v, o);
  fp_load_const(g0, bls::FROB1 + 2 * FPW * L.k);
  fp_load_const(g1, bls::FROB1 + 2 * FPW * L.k + FPW);
  fp_zero(z);
  lz5(n0, z, z, 2, g0, z);
  // (x0 - x1 u)(g0 + g1 u) = (x0 g0 + x1 g1) + (x0 g1 - x1 g0) u
  sel(y0, L.c, g1, g0);
  sel(y1, L.c, n0, g1);
  Wide w;
  w_zero(w);
  w_mac(w, x0, y0);
  w_mac(w, x1, y1);
  Fp r;
  w_reduce(r, w);
  f = r;
}
// x -> x^(p^2): slot k times gamma2[k] in Fp
DEV void frob2(Fp& f, const Lane& L) {
  Fp g, v = f;
  fp_load_const(g, bls::FROB2 + FPW * L.k);
  fp_mul(v, v, g);
  f = v;
}

// Inverse through the norm to Fp6: with g = conj(f) (odd slots negated),
// N = f g = a^2 - b^2 v lies in Fp6 (the even slots 0, 2, 4 hold n0, n1, n2
// of n0 + n1 v + n2 v^2, the odd slots are 0), and f^-1 = g N^-1.  With
// v^3 = xi, N^-1 = (t0 + t1 v + t2 v^2) / d for t0 = n0^2 - xi n1 n2,
// t1 = xi n2^2 - n0 n1, t2 = n1^2 - n0 n2 (lanes (k, 0), k < 3) and
// d = n0 t0 + xi (n2 t1 + n1 t2), d^-1 = conj(d) / (d0^2 + d1^2) (lane (0, 0),
// one Fp inversion).  Two 12-lane products and a few Fp2 operations, with
// values exchanged through the check's LDS area.
// In-line Fp2 products for the inversion below (the out-of-line forms would
// make the caller save its live registers around each call).
DEV void f2_mul(Fp2& r, const Fp2& a, const Fp2& b) {
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
DEV void f2_sqr(Fp2& r, const Fp2& a) {
  Fp sm, df, m;
  fp_add(sm, a.c0, a.c1);
  fp_sub(df, a.c0, a.c1);
  fp_mul(m, a.c0, a.c1);
  fp_mul(r.c0, sm, df);
  fp_dbl(r.c1, m);
}

DEV void inv(Fp& f, const Lane& L) {
  Fp g = f;
  conj(g, L);
  Fp n = f;
  mul_in(n, g, L);  // N
  lds_t* X = L.X;   // N's coefficients: slot s at X + 2 s LE (c0), + LE (c1)
  lds_t* T = L.F;   // t0, t1, t2, d^-1: Fp2 j at T + 2 j LE
  auto ld2 = [&](Fp2& r, const lds_t* q) {
    ld(r.c0, q);
    ld(r.c1, q + LE);
  };
  auto st2 = [&](lds_t* q, const Fp2& v) {
    st(q, v.c0);
    st(q + LE, v.c1);
  };
  __syncthreads();
  if (L.wr) st(X + (2 * L.k + L.c) * LE, n);
  __syncthreads();
  if (L.c == 0 && L.k < 3 && L.wr) {
    // t_k = sq - pr: sq = n_k^2 (times xi for k = 1), pr = the cross product
    // n1 n2 (times xi for k = 0), n0 n1 (k = 1), n0 n2 (k = 2)
    const int sq_slot = L.k == 0 ? 0 : (L.k == 1 ? 4 : 2);
    const int pa = L.k == 0 ? 2 : 0, pb = L.k == 1 ? 2 : 4;
    Fp2 a, b, x, y;
    ld2(a, X + 2 * sq_slot * LE);
    f2_sqr(x, a);
    if (L.k == 1) fp2_mul_xi(x, x);
    ld2(a, X + 2 * pa * LE);
    ld2(b, X + 2 * pb * LE);
    f2_mul(y, a, b);
    if (L.k == 0) fp2_mul_xi(y, y);
    fp2_sub(x, x, y);
    st2(T + 2 * L.k * LE, x);
  }
  __syncthreads();
  if (L.k == 0 && L.c == 0 && L.wr) {
    Fp2 a, b, d, s;
    ld2(a, X + 2 * 2 * LE);  // n1
    ld2(b, T + 2 * 2 * LE);  // t2
    f2_mul(s, a, b);
    ld2(a, X + 2 * 4 * LE);  // n2
    ld2(b, T + 2 * 1 * LE);  // t1
    f2_mul(d, a, b);
    fp2_add(s, s, d);
    fp2_mul_xi(s, s);
    ld2(a, X);              // n0
    ld2(b, T);              // t0
    f2_mul(d, a, b);
    fp2_add(d, d, s);
    Fp nrm, t;
    fp_sqr(nrm, d.c0);
    fp_sqr(t, d.c1);
    fp_add(nrm, nrm, t);
    fp_inv(nrm, nrm);
    fp_mul(d.c0, d.c0, nrm);
    fp_mul(d.c1, d.c1, nrm);
    fp_neg(d.c1, d.c1);
    st2(T + 2 * 3 * LE, d);
  }
  __syncthreads();
  // this lane's coefficient of N^-1: component c of t_{k/2} d^-1 for even k
  Fp h;
  {
    Fp t0, t1, d0, d1, p0, p1;
    const int j = (L.k >> 1) % 3;
    ld(t0, T + 2 * j * LE);
    ld(t1, T + 2 * j * LE + LE);
    ld(d0, T + 2 * 3 * LE);
    ld(d1, T + 2 * 3 * LE + LE);
    // c = 0: t0 d0 - t1 d1, c = 1: t0 d1 + t1 d0
    fp_mul(p0, t0, L.c ? d1 : d0);
    fp_mul(p1, t1, L.c ? d0 : d1);
    Fp q;
    fp_sub(q, p0, p1);
    fp_add(h, p0, p1);
    h = L.c ? h : q;
    if (L.k & 1) fp_zero(h);
  }
  mul_in(g, h, L);
  f = g;
}

// f <- f^x for a cyclotomic element, x = -|x|
DEV void exp_by_x(Fp& f, const Lane& L) {
  const Fp base = f;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    cyc_sqr(f, L);
    if ((BLS_X_ABS >> i) & 1) mul_in(f, base, L);
  }
  conj(f, L);
}

// Same exponent chain as final_exponentiation() in pairing.hpp, written as a
// schedule over five working values that one loop runs: its body holds one
// in-line copy of each operation.  (As straight-line code the chain made 17
// out-of-line calls, each saving and restoring the callee-saved registers
// through scratch memory: most of the kernel's HBM writes.)  Step: R[d] =
// R[s1] * U(R[s2]) (FE_MUL), U(R[s2]) (FE_COPY), R[s1]^-1 (FE_INV) or
// R[s1]^x (FE_EXPX), with U one of id, conj, frob, frob2, cyclotomic square.
// The step index is the same on every lane, so the branches are uniform.
enum { FE_MUL, FE_COPY, FE_INV, FE_EXPX };
enum { FU_ID, FU_CONJ, FU_FROB, FU_FROB2, FU_CYC };
struct FeStep {
  uint8_t op, d, s1, s2, u;
};
// R0 = f (later c), R1 = f1, R2 = t, R3 = a, R4 = b
__constant__ static const FeStep FE_STEPS[] = {
    {FE_INV, 2, 0, 0, FU_ID},      // t = f^-1
    {FE_COPY, 1, 0, 0, FU_CONJ},   // f1 = conj(f)
    {FE_MUL, 1, 1, 2, FU_ID},      // f1 = f^(p^6 - 1)
    {FE_MUL, 1, 1, 1, FU_FROB2},   // f1 = f1^(p^2 + 1)
    {FE_EXPX, 3, 1, 1, FU_ID},     // a = f1^x
    {FE_MUL, 3, 3, 1, FU_CONJ},    // a = f1^(x - 1)
    {FE_EXPX, 2, 3, 3, FU_ID},     // t = a^x
    {FE_MUL, 3, 2, 3, FU_CONJ},    // a = f1^((x - 1)^2)
    {FE_EXPX, 4, 3, 3, FU_ID},     // b = a^x
    {FE_MUL, 4, 4, 3, FU_FROB},    // b = a^(x + p)
    {FE_EXPX, 0, 4, 4, FU_ID},     // c = b^x
    {FE_EXPX, 0, 0, 0, FU_ID},     // c = b^(x^2)
    {FE_MUL, 0, 0, 4, FU_FROB2},   // c = c b^(p^2)
    {FE_MUL, 0, 0, 4, FU_CONJ},    // c = b^(x^2 + p^2 - 1)
    {FE_MUL, 2, 1, 1, FU_CYC},     // t = f1^3
    {FE_MUL, 0, 0, 2, FU_ID},      // c = c f1^3
};
constexpr int FE_N = (int)(sizeof(FE_STEPS) / sizeof(FE_STEPS[0]));

// Selections by masks: a switch over the working values (or a chain of
// selects) is turned into an indexed array in scratch memory.
DEV void fe_get(Fp& x, int s, const Fp& R0, const Fp& R1, const Fp& R2, const Fp& R3, const Fp& R4) {
  const uint32_t m0 = 0u - (uint32_t)(s == 0), m1 = 0u - (uint32_t)(s == 1), m2 = 0u - (uint32_t)(s == 2);
  const uint32_t m3 = 0u - (uint32_t)(s == 3), m4 = 0u - (uint32_t)(s == 4);
#pragma unroll
  for (int k = 0; k < E; k++)
    x.v[k] = (R0.v[k] & m0) | (R1.v[k] & m1) | (R2.v[k] & m2) | (R3.v[k] & m3) | (R4.v[k] & m4);
}
DEV void fe_put(int d, const Fp& x, Fp& R0, Fp& R1, Fp& R2, Fp& R3, Fp& R4) {
  const uint32_t m0 = 0u - (uint32_t)(d == 0), m1 = 0u - (uint32_t)(d == 1), m2 = 0u - (uint32_t)(d == 2);
  const uint32_t m3 = 0u - (uint32_t)(d == 3), m4 = 0u - (uint32_t)(d == 4);
#pragma unroll
  for (int k = 0; k < E; k++) {
    R0.v[k] = (x.v[k] & m0) | (R0.v[k] & ~m0);
    R1.v[k] = (x.v[k] & m1) | (R1.v[k] & ~m1);
    R2.v[k] = (x.v[k] & m2) | (R2.v[k] & ~m2);
    R3.v[k] = (x.v[k] & m3) | (R3.v[k] & ~m3);
    R4.v[k] = (x.v[k] & m4) | (R4.v[k] & ~m4);
  }
}

DEVN void final_exp(Fp& r, const Fp& f, Lane L) {
  Fp R0 = f, R1, R2, R3, R4;
  fp_zero(R1);
  fp_zero(R2);
  fp_zero(R3);
  fp_zero(R4);
#pragma unroll 1
  for (int st = 0; st < FE_N; st++) {
    const FeStep S = FE_STEPS[st];
    Fp x, y;
    fe_get(x, S.s1, R0, R1, R2, R3, R4);
    if (S.op == FE_INV) {
      inv(x, L);
    } else if (S.op == FE_EXPX) {
      exp_by_x(x, L);
    } else {
      fe_get(y, S.s2, R0, R1, R2, R3, R4);
      if (S.u == FU_CONJ)
        conj(y, L);
      else if (S.u == FU_FROB)
        frob(y, L);
      else if (S.u == FU_FROB2)
        frob2(y, L);
      else if (S.u == FU_CYC)
        cyc_sqr(y, L);
      if (S.op == FE_MUL)
        mul_in(x, y, L);
      else
        x = y;
    }
    fe_put(S.d, x, R0, R1, R2, R3, R4);
  }
  r = R0;
}

// Product of the two Miller loops e(P0, Q0) e(P1, Q1) with prepared lines:
// l2 holds the per-check line blocks ([line][q], line_values()), yp_q the y
// of P_q, zero for a skipped pair (its blocks hold the line 1).  Conjugated
// at the end (x < 0).
DEVN void miller(Fp& f, const uint32_t* l2, const Fp& yp0, const Fp& yp1, Lane L) {
  const Fp y0 = yp0, y1 = yp1;
  Fp acc, one, z;
  fp_one(one);
  fp_zero(z);
  sel(acc, L.k == 0 && L.c == 0, one, z);
  int idx = 0;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    if (i != 62) sqr(acc, L);
    const int steps = ((BLS_X_ABS >> i) & 1) ? 2 : 1;
#pragma unroll 1
    for (int s = 0; s < steps; s++) {
#pragma unroll 1
      for (int q = 0; q < 2; q++)
        mul_line(acc, l2 + ((size_t)idx * 2 + q) * LV_N * E, q ? y1 : y0, L);
      idx++;
    }
  }
  conj(acc, L);
  f = acc;
}

// 1 when the element is the identity (the same answer on every lane of the group)
DEVN bool is_one(const Fp& f, Lane L) {
  const Fp v = f;
  __syncthreads();
  Fp one;
  fp_one(one);
  const bool ok = (L.k == 0 && L.c == 0) ? fp_eq(v, one) : fp_is_zero(v);
  if (L.wr) L.X[2 * L.k + L.c] = ok ? 1u : 0u;
  __syncthreads();
  bool all = true;
  for (int i = 0; i < 12; i++) all = all && L.X[i] != 0;
  return all;
}

// Lane setup for a 64-thread block (BLOCK_LDS_WORDS of LDS): group
// g = lane / 12, slot k and component c inside the group.  Lanes 60..63
// (g = 5) run on the last check's region without storing.
DEV Lane lane_setup(uint32_t* lds_all, int lane) {
  Lane L;
  const int g = lane / 12, gl = lane % 12;
  L.wr = g < CHECKS_PER_BLOCK;
  lds_t* base = (lds_t*)lds_all + (L.wr ? g : CHECKS_PER_BLOCK - 1) * LDS_WORDS;
  L.F = base;
  L.X = base + REGION;
  L.Z = (lds_t*)lds_all + CHECKS_PER_BLOCK * LDS_WORDS;
  L.k = gl >> 1;
  L.c = gl & 1;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < E; i++) L.Z[i] = 0;
  }
  return L;
}

}  // namespace p12
