// G1 (E: y^2 = x^3 + 4 over Fp) and G2 (E': y^2 = x^3 + 4(1+u) over Fp2)
// point arithmetic in Jacobian coordinates, ZCash point codec, subgroup
// checks and scalar multiplication (plain windowed and GLV for G1).
//
// Jacobian (X, Y, Z) stands for (X/Z^2, Y/Z^3); Z = 0 is the identity.
#pragma once
#include "field.hpp"

// ------------------------------------------------------------------ field-generic shims
DEV void f_mul(Fp& r, const Fp& a, const Fp& b) { fp_mul(r, a, b); }
DEV void f_sqr(Fp& r, const Fp& a) { fp_sqr(r, a); }
DEV void f_add(Fp& r, const Fp& a, const Fp& b) { fp_add(r, a, b); }
DEV void f_sub(Fp& r, const Fp& a, const Fp& b) { fp_sub(r, a, b); }
DEV void f_dbl(Fp& r, const Fp& a) { fp_dbl(r, a); }
DEV void f_neg(Fp& r, const Fp& a) { fp_neg(r, a); }
DEV bool f_is_zero(const Fp& a) { return fp_is_zero(a); }
DEV bool f_eq(const Fp& a, const Fp& b) { return fp_eq(a, b); }
DEV void f_one(Fp& r) { fp_one(r); }
DEV void f_zero(Fp& r) { fp_zero(r); }
DEVI void f_inv(Fp& r, const Fp& a) { fp_inv(r, a); }
DEVI void f_mul(Fp2& r, const Fp2& a, const Fp2& b) { fp2_mul(r, a, b); }
DEVI void f_sqr(Fp2& r, const Fp2& a) { fp2_sqr(r, a); }
DEV void f_add(Fp2& r, const Fp2& a, const Fp2& b) { fp2_add(r, a, b); }
DEV void f_sub(Fp2& r, const Fp2& a, const Fp2& b) { fp2_sub(r, a, b); }
DEV void f_dbl(Fp2& r, const Fp2& a) { fp2_dbl(r, a); }
DEV void f_neg(Fp2& r, const Fp2& a) { fp2_neg(r, a); }
DEV bool f_is_zero(const Fp2& a) { return fp2_is_zero(a); }
DEV bool f_eq(const Fp2& a, const Fp2& b) { return fp2_eq(a, b); }
DEV void f_one(Fp2& r) { fp2_one(r); }
DEV void f_zero(Fp2& r) { fp2_zero(r); }
DEVI void f_inv(Fp2& r, const Fp2& a) { fp2_inv(r, a); }

template <class F>
struct PtJ { F x, y, z; };
template <class F>
struct PtA { F x, y; };

using G1J = PtJ<Fp>;
using G1A = PtA<Fp>;
using G2J = PtJ<Fp2>;
using G2A = PtA<Fp2>;

template <class F>
DEV bool pt_is_inf(const PtJ<F>& p) { return f_is_zero(p.z); }
template <class F>
DEV void pt_set_inf(PtJ<F>& p) {
  f_one(p.x);
  f_one(p.y);
  f_zero(p.z);
}
template <class F>
DEV void pt_from_affine(PtJ<F>& r, const PtA<F>& a) {
  r.x = a.x;
  r.y = a.y;
  f_one(r.z);
}

// dbl-2009-l (a = 0): 2M + 5S
template <class F>
DEVN void pt_dbl(PtJ<F>& r, const PtJ<F>& p) {
  F A, B, C, D, E, G, t;
  f_sqr(A, p.x);
  f_sqr(B, p.y);
  f_sqr(C, B);
  f_add(t, p.x, B);
  f_sqr(t, t);
  f_sub(t, t, A);
  f_sub(t, t, C);
  f_dbl(D, t);
  f_dbl(E, A);
  f_add(E, E, A);
  f_sqr(G, E);
  F X3, Y3, Z3;
  f_sub(X3, G, D);
  f_sub(X3, X3, D);
  f_mul(Z3, p.y, p.z);
  f_dbl(Z3, Z3);
  f_sub(t, D, X3);
  f_mul(Y3, E, t);
  f_dbl(C, C);
  f_dbl(C, C);
  f_dbl(C, C);
  f_sub(r.y, Y3, C);
  r.x = X3;
  r.z = Z3;  // identity (z = 0) maps to z = 0
}

// add-2007-bl: 11M + 5S, with the exceptional cases handled.
template <class F>
DEVN void pt_add(PtJ<F>& r, const PtJ<F>& p, const PtJ<F>& q) {
  if (pt_is_inf(p)) {
    r = q;
    return;
  }
  if (pt_is_inf(q)) {
    r = p;
    return;
  }
  F Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, rr, V, t;
  f_sqr(Z1Z1, p.z);
  f_sqr(Z2Z2, q.z);
  f_mul(U1, p.x, Z2Z2);
  f_mul(U2, q.x, Z1Z1);
  f_mul(S1, p.y, q.z);
  f_mul(S1, S1, Z2Z2);
  f_mul(S2, q.y, p.z);
  f_mul(S2, S2, Z1Z1);
  f_sub(H, U2, U1);
  f_sub(rr, S2, S1);
  if (f_is_zero(H)) {
    if (f_is_zero(rr)) {
      pt_dbl(r, p);
    } else {
      pt_set_inf(r);
    }
    return;
  }
  f_dbl(I, H);
  f_sqr(I, I);
  f_mul(J, H, I);
  f_dbl(rr, rr);
  f_mul(V, U1, I);
  F X3, Y3, Z3;
  f_sqr(X3, rr);
  f_sub(X3, X3, J);
  f_sub(X3, X3, V);
  f_sub(X3, X3, V);
  f_sub(t, V, X3);
  f_mul(Y3, rr, t);
  f_mul(t, S1, J);
  f_dbl(t, t);
  f_sub(Y3, Y3, t);
  f_add(t, p.z, q.z);
  f_sqr(t, t);
  f_sub(t, t, Z1Z1);
  f_sub(t, t, Z2Z2);
  f_mul(Z3, t, H);
  r.x = X3;
  r.y = Y3;
  r.z = Z3;
}

// madd-2007-bl (q affine, not the identity): 7M + 4S.
template <class F>
DEVN void pt_madd(PtJ<F>& r, const PtJ<F>& p, const PtA<F>& q) {
  if (pt_is_inf(p)) {
    pt_from_affine(r, q);
    return;
  }
  F Z1Z1, U2, S2, H, HH, I, J, rr, V, t;
  f_sqr(Z1Z1, p.z);
  f_mul(U2, q.x, Z1Z1);
  f_mul(S2, q.y, p.z);
  f_mul(S2, S2, Z1Z1);
  f_sub(H, U2, p.x);
  f_sub(rr, S2, p.y);
  if (f_is_zero(H)) {
    if (f_is_zero(rr)) {
      PtJ<F> qj;
      pt_from_affine(qj, q);
      pt_dbl(r, qj);
    } else {
      pt_set_inf(r);
    }
    return;
  }
  f_sqr(HH, H);
  f_dbl(I, HH);
  f_dbl(I, I);
  f_mul(J, H, I);
  f_dbl(rr, rr);
  f_mul(V, p.x, I);
  F X3, Y3, Z3;
  f_sqr(X3, rr);
  f_sub(X3, X3, J);
  f_sub(X3, X3, V);
  f_sub(X3, X3, V);
  f_sub(t, V, X3);
  f_mul(Y3, rr, t);
  f_mul(t, p.y, J);
  f_dbl(t, t);
  f_sub(Y3, Y3, t);
  f_add(t, p.z, H);
  f_sqr(t, t);
  f_sub(t, t, Z1Z1);
  f_sub(Z3, t, HH);
  r.x = X3;
  r.y = Y3;
  r.z = Z3;
}

template <class F>
DEV void pt_neg(PtJ<F>& r, const PtJ<F>& p) {
  r.x = p.x;
  f_neg(r.y, p.y);
  r.z = p.z;
}

// Jacobian -> affine (p must not be the identity)
template <class F>
DEVN void pt_to_affine(PtA<F>& r, const PtJ<F>& p) {
  F zi, zi2, zi3;
  f_inv(zi, p.z);
  f_sqr(zi2, zi);
  f_mul(zi3, zi2, zi);
  f_mul(r.x, p.x, zi2);
  f_mul(r.y, p.y, zi3);
}

// P == (x, y) affine?  (p may be the identity -> false)
template <class F>
DEVN bool pt_eq_affine(const PtJ<F>& p, const PtA<F>& q) {
  if (pt_is_inf(p)) return false;
  F z2, z3, t;
  f_sqr(z2, p.z);
  f_mul(t, q.x, z2);
  if (!f_eq(t, p.x)) return false;
  f_mul(z3, z2, p.z);
  f_mul(t, q.y, z3);
  return f_eq(t, p.y);
}

// r = [k] p, k a small (<= 64 bit) constant, left-to-right
template <class F>
DEVN void pt_mul_u64(PtJ<F>& r, const PtJ<F>& p, uint64_t k) {
  PtJ<F> acc;
  pt_set_inf(acc);
#pragma unroll 1
  for (int i = 63; i >= 0; i--) {
    pt_dbl(acc, acc);
    if ((k >> i) & 1) pt_add(acc, acc, p);
  }
  r = acc;
}

// r = [k] q for an affine base, k a non-zero small constant: starts at the top
// set bit, mixed additions.
template <class F>
DEVN void pt_mul_u64_affine(PtJ<F>& r, const PtA<F>& q, uint64_t k) {
  PtJ<F> acc;
  pt_from_affine(acc, q);
  const int top = 63 - __builtin_clzll(k);
#pragma unroll 1
  for (int i = top - 1; i >= 0; i--) {
    pt_dbl(acc, acc);
    if ((k >> i) & 1) pt_madd(acc, acc, q);
  }
  r = acc;
}

// r = [k] q for a canonical scalar k given as N little-endian 32-bit limbs
// and an affine base (binary left-to-right with mixed additions).
template <class F, int N>
DEVN void pt_mul_limbs_affine(PtJ<F>& r, const PtA<F>& q, const uint32_t* k) {
  PtJ<F> acc;
  pt_set_inf(acc);
#pragma unroll 1
  for (int i = N * 32 - 1; i >= 0; i--) {
    pt_dbl(acc, acc);
    if ((k[i >> 5] >> (i & 31)) & 1) pt_madd(acc, acc, q);
  }
  r = acc;
}

// ------------------------------------------------------------------ inlined G1 group law
// The same formulas as pt_dbl / pt_madd / pt_add (dbl-2009-l, madd-2007-bl,
// add-2007-bl) for G1, expanded into the calling loop instead of called (no
// call frame in scratch memory), with the operations ordered so that at most
// six field temporaries are live.  The rare exceptional cases (doubling
// inside an addition) call the out-of-line forms.

// k p in normalized 28-bit limbs (k p < 2^392)
template <uint32_t K>
struct PMul {
  uint32_t v[14];
  constexpr PMul() : v() {
    uint64_t c = 0;
    for (int i = 0; i < 14; i++) {
      c += (uint64_t)K * bls::P28[i];
      v[i] = (uint32_t)(c & 0x0fffffffu);
      c >>= 28;
    }
  }
};
static constexpr PMul<6> P6_L{};
static constexpr PMul<16> P16_L{};
static constexpr PMul<24> P24_L{};
// a value below 2^392 in normalized limbs (r.v[k] < 2^28) that is below
// (K + 1) 2p for K < 200: minus q 2p, q = floor(top limb / (top limb of 2p +
// 1)) (never above the true quotient, at most one below it), then one
// conditional subtraction of 2p: the result is below 2p.
DEV void fp_reduce_multiple(Fp& r) {
  constexpr uint32_t T = bls::P2_28[13] + 1;
  const uint32_t q = r.v[13] / T;
  Fp d, e;
  int64_t b1 = 0;
  int32_t b2 = 0;
#pragma unroll
  for (int i = 0; i < 14; i++) {
    // q 2p_i reaches 2^36: 64-bit
    const int64_t u = (int64_t)r.v[i] - (int64_t)q * bls::P2_28[i] + b1;
    b1 = u >> 28;
    d.v[i] = (uint32_t)u & FP_MASK;
    const int32_t t = (int32_t)d.v[i] - (int32_t)bls::P2_28[i] + b2;
    b2 = t >> 28;
    e.v[i] = (uint32_t)t & FP_MASK;
  }
#pragma unroll
  for (int i = 0; i < 14; i++) r.v[i] = b2 < 0 ? d.v[i] : e.v[i];
}

// p <- 2 p, dbl-2009-l with lazy sums: the terms that only enter products
// stay unnormalized (X + B, 3A) or are formed in one signed limb pass, and
// each output is one signed limb pass plus fp_reduce_multiple.  Bounds (p's
// coordinates below 2p): X + B below 4p with limbs below 2^29; E = 3A below
// 6p with limbs below 3 2^28 (E^2 < 36 p^2 < p R; every product column stays
// below 2^64); D' = T + 4p - A - C in (0, 6p), D = 2 D'; X3 = F + 24p - 4 D'
// in (0, 26p); W = 2 D' + 2p - X3 in (0, 14p); Y3 = E W + 16p - 8 C in
// (0, 18p).
DEV void g1_dbl_in(G1J& p) {
  Fp A, B, C, D, E, t;
  fp_sqr(A, p.x);  // A = X^2
  fp_sqr(B, p.y);  // B = Y^2
  fp_mul(t, p.y, p.z);
  fp_dbl(p.z, t);  // Z3 = 2 Y Z (the identity keeps Z = 0)
  fp_sqr(C, B);    // C = B^2
#pragma unroll
  for (int i = 0; i < 14; i++) t.v[i] = p.x.v[i] + B.v[i];  // X + B
  fp_sqr(t, t);
  {
    int32_t cy = 0;  // D' = (X + B)^2 + 4p - A - C
#pragma unroll
    for (int i = 0; i < 14; i++) {
      const int32_t u = (int32_t)(t.v[i] + bls::P4_28[i]) - (int32_t)A.v[i] - (int32_t)C.v[i] + cy;
      cy = u >> 28;
      D.v[i] = (uint32_t)u & FP_MASK;
    }
  }
#pragma unroll
  for (int i = 0; i < 14; i++) E.v[i] = 3u * A.v[i];  // E = 3A
  fp_sqr(t, E);                                        // F = E^2
  {
    int32_t cy = 0;  // X3 = F + 24p - 4 D'
#pragma unroll
    for (int i = 0; i < 14; i++) {
      const int32_t u = (int32_t)(t.v[i] + P24_L.v[i]) - (int32_t)(D.v[i] << 2) + cy;
      cy = u >> 28;
      p.x.v[i] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.x);
  {
    int32_t cy = 0;  // W = 2 D' + 2p - X3
#pragma unroll
    for (int i = 0; i < 14; i++) {
      const int32_t u = (int32_t)((D.v[i] << 1) + bls::P2_28[i]) - (int32_t)p.x.v[i] + cy;
      cy = u >> 28;
      D.v[i] = (uint32_t)u & FP_MASK;
    }
  }
  fp_mul(t, E, D);  // E W
  {
    int32_t cy = 0;  // Y3 = E W + 16p - 8C
#pragma unroll
    for (int i = 0; i < 14; i++) {
      const int32_t u = (int32_t)(t.v[i] + P16_L.v[i]) - (int32_t)(C.v[i] << 3) + cy;
      cy = u >> 28;
      p.y.v[i] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.y);
}

// p <- p + q, q affine and not the identity (madd-2007-bl with lazy sums:
// r = 2 (S2 - Y1) and I = 4 HH stay unnormalized products operands, limbs
// below 2^29 and 2^30; Z3 = 2 Z1 H; X3 = r^2 + 6p - J - 2V in (0, 8p) and
// Y3 = r W + 4p - 2 Y1 J in (0, 6p), W = V + 2p - X3 in (0, 4p), each one
// signed limb pass, the outputs reduced below 2p)
DEV void g1_madd_in(G1J& p, const G1A& q) {
  if (fp_is_zero(p.z)) {
    pt_from_affine(p, q);
    return;
  }
  Fp t0, t1, t2, t3, t4;
  fp_sqr(t0, p.z);          // Z1Z1
  fp_mul(t1, q.x, t0);      // U2
  fp_mul(t2, q.y, p.z);
  fp_mul(t2, t2, t0);       // S2
  fp_sub(t1, t1, p.x);      // H = U2 - X1
  fp_sub(t2, t2, p.y);      // S2 - Y1
  if (fp_is_zero(t1)) {
    if (fp_is_zero(t2)) {
      // the out-of-line doubling works on locals only: passing p itself would
      // take its address and keep the caller's accumulator in scratch memory
      // (an inline doubling here instead doubles the code of every caller and
      // slowed the single-lane loops)
      G1J qj, r;
      pt_from_affine(qj, q);
      pt_dbl(r, qj);
      p = r;
    } else {
      pt_set_inf(p);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < 14; k++) t2.v[k] <<= 1;  // r = 2 (S2 - Y1), lazy
  fp_mul(p.z, p.z, t1);     // Z3 = 2 Z1 H
  fp_dbl(p.z, p.z);
  fp_sqr(t3, t1);           // HH
#pragma unroll
  for (int k = 0; k < 14; k++) t3.v[k] <<= 2;  // I = 4 HH, lazy
  fp_mul(t0, t1, t3);       // J = H I
  fp_mul(t1, p.x, t3);      // V = X1 I
  fp_mul(t4, p.y, t0);      // Y1 J
  fp_sqr(t3, t2);           // r^2
  {
    int32_t cy = 0;  // X3 = r^2 + 6p - J - 2V
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t3.v[k] + P6_L.v[k]) - (int32_t)t0.v[k] - (int32_t)(t1.v[k] << 1) + cy;
      cy = u >> 28;
      p.x.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.x);
  {
    int32_t cy = 0;  // W = V + 2p - X3
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t1.v[k] + bls::P2_28[k]) - (int32_t)p.x.v[k] + cy;
      cy = u >> 28;
      t1.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_mul(t3, t2, t1);       // r W
  {
    int32_t cy = 0;  // Y3 = r W + 4p - 2 Y1 J
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t3.v[k] + bls::P4_28[k]) - (int32_t)(t4.v[k] << 1) + cy;
      cy = u >> 28;
      p.y.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.y);
}

// p <- p + q (Jacobian; q must not alias p); add-2007-bl with lazy sums:
// Z3 = ((Z1 + Z2)^2 + 4p - Z1Z1 - Z2Z2) H, r = 2 (S2 - S1) unnormalized,
// X3 = r^2 + 6p - J - 2V and Y3 = r W + 4p - 2 S1 J (W = V + 2p - X3), each
// one signed limb pass, the outputs reduced below 2p
DEV void g1_add_in(G1J& p, const G1J& q) {
  if (fp_is_zero(q.z)) return;
  if (fp_is_zero(p.z)) {
    p = q;
    return;
  }
  Fp t0, t1, t2, t3, t4, t5;
  fp_sqr(t0, p.z);          // Z1Z1
  fp_sqr(t1, q.z);          // Z2Z2
  fp_mul(t2, p.x, t1);      // U1
  fp_mul(t3, q.x, t0);      // U2
  fp_mul(t4, p.y, q.z);
  fp_mul(t4, t4, t1);       // S1
  fp_mul(t5, q.y, p.z);
  fp_mul(t5, t5, t0);       // S2
  fp_sub(t3, t3, t2);       // H = U2 - U1
  fp_sub(t5, t5, t4);       // S2 - S1
  if (fp_is_zero(t3)) {
    if (fp_is_zero(t5)) {
      // locals only, as in g1_madd_in; z is normalized here so that `a` is
      // not a plain copy, which the compiler would replace by p itself and
      // so keep the caller's accumulator in scratch memory on every addition
      G1J a, r;
      Fp zero;
      fp_zero(zero);
      a.x = p.x;
      a.y = p.y;
      fp_add(a.z, p.z, zero);
      pt_dbl(r, a);
      p = r;
    } else {
      pt_set_inf(p);
    }
    return;
  }
  {
    Fp s;
#pragma unroll
    for (int k = 0; k < 14; k++) s.v[k] = p.z.v[k] + q.z.v[k];  // Z1 + Z2, lazy
    fp_sqr(s, s);
    int32_t cy = 0;  // (Z1 + Z2)^2 + 4p - Z1Z1 - Z2Z2, in (0, 6p)
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(s.v[k] + bls::P4_28[k]) - (int32_t)t0.v[k] - (int32_t)t1.v[k] + cy;
      cy = u >> 28;
      s.v[k] = (uint32_t)u & FP_MASK;
    }
    fp_mul(p.z, s, t3);     // Z3
  }
#pragma unroll
  for (int k = 0; k < 14; k++) t5.v[k] <<= 1;  // r = 2 (S2 - S1), lazy
#pragma unroll
  for (int k = 0; k < 14; k++) t0.v[k] = t3.v[k] << 1;  // 2H, lazy
  fp_sqr(t0, t0);           // I = (2H)^2
  fp_mul(t1, t3, t0);       // J = H I
  fp_mul(t3, t2, t0);       // V = U1 I
  fp_mul(t4, t4, t1);       // S1 J
  fp_sqr(t0, t5);           // r^2
  {
    int32_t cy = 0;  // X3 = r^2 + 6p - J - 2V
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t0.v[k] + P6_L.v[k]) - (int32_t)t1.v[k] - (int32_t)(t3.v[k] << 1) + cy;
      cy = u >> 28;
      p.x.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.x);
  {
    int32_t cy = 0;  // W = V + 2p - X3
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t3.v[k] + bls::P2_28[k]) - (int32_t)p.x.v[k] + cy;
      cy = u >> 28;
      t3.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_mul(t0, t5, t3);       // r W
  {
    int32_t cy = 0;  // Y3 = r W + 4p - 2 S1 J
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t0.v[k] + bls::P4_28[k]) - (int32_t)(t4.v[k] << 1) + cy;
      cy = u >> 28;
      p.y.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.y);
}

// ------------------------------------------------------------------ G1 in XYZZ coordinates
// (X, Y, ZZ, ZZZ) stands for (X/ZZ, Y/ZZZ) with ZZ^3 = ZZZ^2; ZZ = 0 is the
// identity.  A mixed addition (madd-2008-s) costs 8M + 2S and 7 additions,
// against 7M + 4S and 13 for madd-2007-bl in Jacobian coordinates, so long
// sums of affine terms (the fixed-base comb) accumulate in this form and
// convert to Jacobian once.
struct G1Z {
  Fp x, y, zz, zzz;
};
DEV void g1z_set_inf(G1Z& p) {
  fp_one(p.x);
  fp_one(p.y);
  fp_zero(p.zz);
  fp_zero(p.zzz);
}
// 4p with every limb but the top one at least 2^28 - 1 (no borrows when a
// normalized value below 2p is subtracted from it limb by limb)
struct P4Redundant {
  uint32_t v[14];
  constexpr P4Redundant() : v() {
    for (int i = 0; i < 14; i++) v[i] = bls::P4_28[i];
    v[0] += 1u << 28;
    for (int i = 1; i < 13; i++) v[i] += (1u << 28) - 1;
    v[13] -= 1;
  }
};
static constexpr P4Redundant P4_RED{};
// y <- 4p - y when neg (value below 4p, limbs below 2^29: an operand for a
// product only), else y
DEV void fp_cond_neg_lazy(Fp& y, bool neg) {
#pragma unroll
  for (int k = 0; k < 14; k++) y.v[k] = neg ? P4_RED.v[k] - y.v[k] : y.v[k];
}
// 2 Q for an affine Q (mdbl-2008-s-1), the rare case of a mixed addition of
// equal points; Q.y may be lazily negated
DEVN void g1z_dbl_affine(G1Z& r, const G1A& q) {
  Fp y, z, U, V, W, S, M, t;
  fp_zero(z);
  fp_add(y, q.y, z);  // below 2p, normalized
  fp_dbl(U, y);
  fp_sqr(V, U);
  fp_mul(W, U, V);
  fp_mul(S, q.x, V);
  fp_sqr(t, q.x);
  fp_dbl(M, t);
  fp_add(M, M, t);
  fp_sqr(t, M);
  fp_sub(t, t, S);
  fp_sub(r.x, t, S);
  fp_sub(t, S, r.x);
  fp_mul(t, M, t);
  fp_mul(S, W, y);
  fp_sub(r.y, t, S);
  r.zz = V;
  r.zzz = W;
}
// p <- p + q, q affine and not the identity (q.y may be lazily negated:
// it only enters products and the normalizing additions below)
DEV void g1z_madd(G1Z& p, const G1A& q) {
  Fp z;
  fp_zero(z);
  if (fp_is_zero(p.zz)) {
    p.x = q.x;
    fp_add(p.y, q.y, z);
    fp_one(p.zz);
    fp_one(p.zzz);
    return;
  }
  Fp P, R, PP, PPP, Q, t;
  fp_mul(P, q.x, p.zz);   // U2
  fp_sub(P, P, p.x);      // P = U2 - X1
  fp_mul(R, q.y, p.zzz);  // S2
  fp_sub(R, R, p.y);      // R = S2 - Y1
  if (fp_is_zero(P)) {
    if (fp_is_zero(R)) {
      // a new point for the out-of-line doubling (y normalized here, so it is
      // not a plain copy the compiler would replace by q): passing q itself
      // would take the address of the caller's point, which then goes to
      // scratch memory on every addition, not only on this path
      G1A qa;
      qa.x = q.x;
      fp_add(qa.y, q.y, z);
      G1Z d;
      g1z_dbl_affine(d, qa);
      p = d;
    } else {
      g1z_set_inf(p);
    }
    return;
  }
  fp_sqr(PP, P);
  fp_mul(PPP, P, PP);
  fp_mul(Q, p.x, PP);
  fp_mul(p.zz, p.zz, PP);
  fp_mul(p.zzz, p.zzz, PPP);
  fp_sqr(t, R);
  {
    int32_t cy = 0;  // X3 = R^2 + 6p - PPP - 2Q, in (0, 8p)
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(t.v[k] + P6_L.v[k]) - (int32_t)PPP.v[k] - (int32_t)(Q.v[k] << 1) + cy;
      cy = u >> 28;
      p.x.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_reduce_multiple(p.x);
  {
    int32_t cy = 0;  // W = Q + 2p - X3, in (0, 4p)
#pragma unroll
    for (int k = 0; k < 14; k++) {
      const int32_t u = (int32_t)(Q.v[k] + bls::P2_28[k]) - (int32_t)p.x.v[k] + cy;
      cy = u >> 28;
      t.v[k] = (uint32_t)u & FP_MASK;
    }
  }
  fp_mul(t, R, t);        // Y3 = R W - Y1 PPP
  fp_mul(PP, p.y, PPP);
  fp_sub(p.y, t, PP);
}
// The same point in Jacobian coordinates: (X ZZ^2, Y ZZ^3, ZZZ)
DEV void g1z_to_jacobian(G1J& r, const G1Z& p) {
  Fp z2, z3;
  fp_sqr(z2, p.zz);
  fp_mul(z3, z2, p.zz);
  fp_mul(r.x, p.x, z2);
  fp_mul(r.y, p.y, z3);
  r.z = p.zzz;
}

// ------------------------------------------------------------------ G1 specifics
DEV void g1_load_affine_const(G1A& r, const uint32_t* c) {
  fp_load_const(r.x, c);
  fp_load_const(r.y, c + FPW);
}

// phi(x, y) = (beta x, y) acts on G1 as multiplication by lambda = x^2 - 1.
DEV void g1_endo(G1J& r, const G1J& p) {
  Fp beta;
  fp_load_const(beta, bls::GLV_BETA);
  fp_mul(r.x, p.x, beta);
  r.y = p.y;
  r.z = p.z;
}

// Subgroup test for points of E(Fp) (Bowe, "Faster subgroup checks for
// BLS12-381", 2019): P in G1 iff phi'(P) = [-x^2] P where phi' = phi^2 acts
// as lambda^2 = -x^2.  Here: [x^2] P == (beta^2 x, -y).
// The point is given in affine form (not the identity).
//
// |x| = 0xd201000000010000 has the bits 63, 62, 60, 57, 48 and 16 set: after
// the top bit, runs of 1, 2, 3, 9 and 32 doublings each end in an addition
// of the base and 16 doublings follow.  Both products [|x|] P and
// [|x|] ([|x|] P) walk these runs with the in-line group law on locals (no
// call frames): 126 doublings, 5 mixed and 5 full additions.
DEV int g1_x_run(int s) { return s == 0 ? 1 : s == 1 ? 2 : s == 2 ? 3 : s == 3 ? 9 : s == 4 ? 32 : 16; }
// `load(a)` fetches the affine point (not the identity) wherever it is
// needed, so that callers holding it in memory keep it out of registers
// across the doubling runs.
template <class Load>
DEV bool g1_subgroup_test(Load load) {
  G1J t;
  {
    G1A pa;
    load(pa);
    pt_from_affine(t, pa);
  }
#pragma unroll 1
  for (int s = 0; s < 6; s++) {
#pragma unroll 1
    for (int i = 0; i < g1_x_run(s); i++) g1_dbl_in(t);
    if (s < 5) {
      G1A pa;
      load(pa);
      g1_madd_in(t, pa);
    }
  }
  const G1J q = t;
#pragma unroll 1
  for (int s = 0; s < 6; s++) {
#pragma unroll 1
    for (int i = 0; i < g1_x_run(s); i++) g1_dbl_in(t);
    if (s < 5) g1_add_in(t, q);
  }
  // t == (beta^2 x, -y)?  (t the identity: no, as the point is not)
  if (fp_is_zero(t.z)) return false;
  G1A pa;
  load(pa);
  Fp beta, u, z2;
  fp_load_const(beta, bls::GLV_BETA);
  fp_mul(u, pa.x, beta);
  fp_mul(u, u, beta);
  fp_sqr(z2, t.z);
  fp_mul(u, u, z2);
  if (!fp_eq(u, t.x)) return false;
  fp_mul(z2, z2, t.z);
  fp_neg(u, pa.y);
  fp_mul(u, u, z2);
  return fp_eq(u, t.y);
}
DEV bool g1_in_subgroup(const G1A& pa) {
  return g1_subgroup_test([&](G1A& a) { a = pa; });
}
// The point stored affine at pm (G1A_WORDS words).
DEV bool g1_in_subgroup_at(const uint32_t* pm) {
  return g1_subgroup_test([&](G1A& a) {
#pragma unroll
    for (int k = 0; k < FPW; k++) {
      a.x.v[k] = pm[k];
      a.y.v[k] = pm[FPW + k];
    }
  });
}

// Decode a 48-byte compressed G1 point (bls12_381 G1Affine::from_compressed
// semantics): compression flag set; infinity encoding must be 0xc0 00..00;
// x < p; on curve; in the prime-order subgroup.  Returns false on any failure.
DEVN bool g1_decompress(G1J& r, const uint8_t* in, bool check_subgroup = true) {
  uint8_t b[48];
#pragma unroll
  for (int i = 0; i < 48; i++) b[i] = in[i];
  const uint32_t flags = b[0];
  const bool comp = (flags >> 7) & 1, inf = (flags >> 6) & 1, sign = (flags >> 5) & 1;
  if (!comp) return false;
  b[0] &= 0x1f;
  if (inf) {
    uint32_t o = sign;
#pragma unroll
    for (int i = 0; i < 48; i++) o |= b[i];
    pt_set_inf(r);
    return o == 0;
  }
  Fp x, y, rhs, bb;
  if (!fp_from_be_bytes(x, b)) return false;
  fp_to_mont(x, x);
  fp_sqr(rhs, x);
  fp_mul(rhs, rhs, x);
  fp_load_const(bb, bls::FP_B);
  fp_add(rhs, rhs, bb);
  if (!fp_sqrt(y, rhs)) return false;
  if (fp_lex_largest(y) != sign) fp_neg(y, y);
  r.x = x;
  r.y = y;
  fp_one(r.z);
  if (!check_subgroup) return true;
  G1A a;
  a.x = x;
  a.y = y;
  return g1_in_subgroup(a);
}

DEVN void g1_compress_affine(uint8_t* out, const G1A& a) {
  Fp c;
  fp_from_mont(c, a.x);
  fp_to_be_bytes(out, c);
  uint8_t f = 0x80;
  if (fp_lex_largest(a.y)) f |= 0x20;
  out[0] |= f;
}
DEVN void g1_compress(uint8_t* out, const G1J& p) {
  if (pt_is_inf(p)) {
    out[0] = 0xc0;
    for (int i = 1; i < 48; i++) out[i] = 0;
    return;
  }
  G1A a;
  pt_to_affine(a, p);
  g1_compress_affine(out, a);
}

// ------------------------------------------------------------------ G2 specifics
DEV void g2_load_affine_const(G2A& r, const uint32_t* c) {
  fp_load_const(r.x.c0, c);
  fp_load_const(r.x.c1, c + FPW);
  fp_load_const(r.y.c0, c + 2 * FPW);
  fp_load_const(r.y.c1, c + 3 * FPW);
}

DEVN bool g2_in_subgroup(const G2J& q) {
  if (pt_is_inf(q)) return true;
  G2A a;
  pt_to_affine(a, q);
  G2J t;
  pt_mul_limbs_affine<Fp2, 8>(t, a, bls::R_MOD);
  return pt_is_inf(t);
}

// 96-byte compressed G2 (x.c1 || x.c0, flags in the first byte).
DEVN bool g2_decompress(G2J& r, const uint8_t* in) {
  uint8_t b[96];
  for (int i = 0; i < 96; i++) b[i] = in[i];
  const uint32_t flags = b[0];
  const bool comp = (flags >> 7) & 1, inf = (flags >> 6) & 1, sign = (flags >> 5) & 1;
  if (!comp) return false;
  b[0] &= 0x1f;
  if (inf) {
    uint32_t o = sign;
    for (int i = 0; i < 96; i++) o |= b[i];
    pt_set_inf(r);
    return o == 0;
  }
  Fp2 x, y, rhs, bb;
  if (!fp_from_be_bytes(x.c1, b) || !fp_from_be_bytes(x.c0, b + 48)) return false;
  fp_to_mont(x.c0, x.c0);
  fp_to_mont(x.c1, x.c1);
  fp2_sqr(rhs, x);
  fp2_mul(rhs, rhs, x);
  fp_load_const(bb.c0, bls::FP_B);
  bb.c1 = bb.c0;
  fp2_add(rhs, rhs, bb);
  if (!fp2_sqrt(y, rhs)) return false;
  if (fp2_lex_largest(y) != sign) fp2_neg(y, y);
  r.x = x;
  r.y = y;
  fp2_one(r.z);
  return g2_in_subgroup(r);
}

DEVN void g2_compress(uint8_t* out, const G2J& p) {
  if (pt_is_inf(p)) {
    out[0] = 0xc0;
    for (int i = 1; i < 96; i++) out[i] = 0;
    return;
  }
  G2A a;
  pt_to_affine(a, p);
  Fp c;
  fp_from_mont(c, a.x.c1);
  fp_to_be_bytes(out, c);
  fp_from_mont(c, a.x.c0);
  fp_to_be_bytes(out + 48, c);
  uint8_t f = 0x80;
  if (fp2_lex_largest(a.y)) f |= 0x20;
  out[0] |= f;
}

// ------------------------------------------------------------------ GLV scalar split
// k = k1 + k2 * lambda with lambda = x^2 - 1 (127.4 bits): k2 = floor(k / lambda),
// k1 = k mod lambda; both fit in 128 bits and are non-negative.
DEVN void glv_split(uint32_t k1[4], uint32_t k2[4], const uint32_t k[8]) {
  // q = floor(k * mu / 2^256), mu = floor(2^256 / lambda) (129 bits) ; q <= true quotient
  uint32_t prod[17];
#pragma unroll
  for (int i = 0; i < 17; i++) prod[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      c += (uint64_t)k[i] * bls::GLV_MU[j] + prod[i + j];
      prod[i + j] = (uint32_t)c;
      c >>= 32;
    }
    prod[i + 8] = (uint32_t)c;
  }
  uint32_t q[5];
#pragma unroll
  for (int i = 0; i < 5; i++) q[i] = prod[8 + i];
  // rem = k - q * lambda
  uint32_t ql[9];
#pragma unroll
  for (int i = 0; i < 9; i++) ql[i] = 0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (i + j < 9) {
        c += (uint64_t)q[i] * bls::GLV_LAMBDA[j] + ql[i + j];
        ql[i + j] = (uint32_t)c;
        c >>= 32;
      }
    }
    if (i + 4 < 9) ql[i + 4] = (uint32_t)c;
  }
  uint32_t rem[8];
  uint32_t borrow = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    uint64_t x = (uint64_t)k[i] - ql[i] - borrow;
    rem[i] = (uint32_t)x;
    borrow = (uint32_t)(x >> 63);
  }
  // at most a few corrections: while rem >= lambda
#pragma unroll 1
  for (int it = 0; it < 4; it++) {
    uint32_t d[8];
    uint32_t b2 = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint64_t x = (uint64_t)rem[i] - (i < 4 ? bls::GLV_LAMBDA[i] : 0u) - b2;
      d[i] = (uint32_t)x;
      b2 = (uint32_t)(x >> 63);
    }
    if (b2) break;
#pragma unroll
    for (int i = 0; i < 8; i++) rem[i] = d[i];
    uint32_t c = 1;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      uint64_t x = (uint64_t)q[i] + c;
      q[i] = (uint32_t)x;
      c = (uint32_t)(x >> 32);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    k1[i] = rem[i];
    k2[i] = q[i];
  }
}

// [k] P on G1 (P affine, not the identity) with the GLV split
// k = k1 + k2 lambda: 128 shared doublings, mixed additions of P and phi(P).
DEVN void g1_mul_glv(G1J& r, const G1A& p, const uint32_t k[8]) {
  uint32_t k1[4], k2[4];
  glv_split(k1, k2, k);
  G1A pe;
  Fp beta;
  fp_load_const(beta, bls::GLV_BETA);
  fp_mul(pe.x, p.x, beta);
  pe.y = p.y;
  G1J acc;
  pt_set_inf(acc);
#pragma unroll 1
  for (int i = 127; i >= 0; i--) {
    pt_dbl(acc, acc);
    if ((k1[i >> 5] >> (i & 31)) & 1) pt_madd(acc, acc, p);
    if ((k2[i >> 5] >> (i & 31)) & 1) pt_madd(acc, acc, pe);
  }
  r = acc;
}
