// Pairing check spread over 6 lanes.
//
// An Fp12 element is held as its six coefficients in the basis
// 1, w, w^2, .., w^5 (w^6 = xi = 1 + u): lane k of an 8-lane group owns the
// Fp2 coefficient of w^k (lanes 6 and 7 idle).  Products are formed slot by
// slot from the other lanes' coefficients, exchanged through LDS:
//   (a b)_k = sum_i a_i b_{k-i}  (times xi when i > k)
// so a full multiplication is 6 Fp2 products per lane (18 Fp multiplications
// of latency instead of 54), a squaring 4, a sparse line 2 + 1.  The
// cyclotomic squaring (Granger-Scott) runs its three Fp4 squarings on lanes
// 0..2.  Slot k maps to the tower coordinates as c_{k mod 2}.c_{k div 2}.
//
// Every function here contains __syncthreads(): all 64 threads of the block
// (8 groups, one wavefront) must call it in the same order.  Blocks are
// exactly 64 threads.
#pragma once
#include "pairing.hpp"

#define G6_LANES 8
// words per Fp2 slot in LDS
#define G6_F2 (2 * FPW)
// LDS words per group: two Fp12 operands of 6 Fp2 slots, plus flags
#define G6_LDS_WORDS (2 * 6 * 2 * FPW + 8)

DEV void g6_put(uint32_t* dst, const Fp2& v) {
#pragma unroll
  for (int i = 0; i < FPW; i++) {
    dst[i] = v.c0.v[i];
    dst[FPW + i] = v.c1.v[i];
  }
}
DEV void g6_get(Fp2& v, const uint32_t* src) {
#pragma unroll
  for (int i = 0; i < FPW; i++) {
    v.c0.v[i] = src[i];
    v.c1.v[i] = src[FPW + i];
  }
}

// f <- f * g
DEVN void g6_mul(Fp2& f, const Fp2& g, int k, uint32_t* lds) {
  uint32_t* A = lds;
  uint32_t* B = lds + 6 * G6_F2;
  __syncthreads();
  if (k < 6) {
    g6_put(A + G6_F2 * k, f);
    g6_put(B + G6_F2 * k, g);
  }
  __syncthreads();
  if (k < 6) {
    Fp2 acc, x, y, p;
    fp2_zero(acc);
    for (int i = 0; i < 6; i++) {
      const int j = (k - i + 6) % 6;
      g6_get(x, A + G6_F2 * i);
      g6_get(y, B + G6_F2 * j);
      fp2_mul(p, x, y);
      if (i > k) fp2_mul_xi(p, p);
      fp2_add(acc, acc, p);
    }
    f = acc;
  }
}

// f <- f^2 (symmetric terms counted once and doubled)
__constant__ static const int8_t G6_SQ_TERMS[6][4][4] = {
    // {i, j, times2, times_xi}; i < 0 marks an empty term
    {{0, 0, 0, 0}, {1, 5, 1, 1}, {2, 4, 1, 1}, {3, 3, 0, 1}},
    {{0, 1, 1, 0}, {2, 5, 1, 1}, {3, 4, 1, 1}, {-1, 0, 0, 0}},
    {{0, 2, 1, 0}, {1, 1, 0, 0}, {3, 5, 1, 1}, {4, 4, 0, 1}},
    {{0, 3, 1, 0}, {1, 2, 1, 0}, {4, 5, 1, 1}, {-1, 0, 0, 0}},
    {{0, 4, 1, 0}, {1, 3, 1, 0}, {2, 2, 0, 0}, {5, 5, 0, 1}},
    {{0, 5, 1, 0}, {1, 4, 1, 0}, {2, 3, 1, 0}, {-1, 0, 0, 0}}};

DEVN void g6_sqr(Fp2& f, int k, uint32_t* lds) {
  uint32_t* A = lds;
  __syncthreads();
  if (k < 6) g6_put(A + G6_F2 * k, f);
  __syncthreads();
  if (k < 6) {
    Fp2 acc, x, y, p;
    fp2_zero(acc);
    for (int t = 0; t < 4; t++) {
      const int i = G6_SQ_TERMS[k][t][0];
      if (i < 0) continue;
      const int j = G6_SQ_TERMS[k][t][1];
      g6_get(x, A + G6_F2 * i);
      g6_get(y, A + G6_F2 * j);
      fp2_mul(p, x, y);
      if (G6_SQ_TERMS[k][t][2]) fp2_dbl(p, p);
      if (G6_SQ_TERMS[k][t][3]) fp2_mul_xi(p, p);
      fp2_add(acc, acc, p);
    }
    f = acc;
  }
}

// f <- f * (l0 + l2 w^2 + l3 w^3), l3 in Fp (a Miller-loop line)
DEVN void g6_mul_line(Fp2& f, const Fp2& l0, const Fp2& l2, const Fp& l3, int k, uint32_t* lds) {
  uint32_t* A = lds;
  __syncthreads();
  if (k < 6) g6_put(A + G6_F2 * k, f);
  __syncthreads();
  if (k < 6) {
    Fp2 acc, x, p;
    fp2_mul(acc, f, l0);
    g6_get(x, A + G6_F2 * ((k + 4) % 6));
    fp2_mul(p, x, l2);
    if (k < 2) fp2_mul_xi(p, p);
    fp2_add(acc, acc, p);
    g6_get(x, A + G6_F2 * ((k + 3) % 6));
    fp2_mul_fp(p, x, l3);
    if (k < 3) fp2_mul_xi(p, p);
    fp2_add(f, acc, p);
  }
}

DEV void g6_conj(Fp2& f, int k) {
  if (k & 1) fp2_neg(f, f);
}

// x -> x^p: slot k is conjugated and scaled by xi^(k (p-1)/6)
DEVN void g6_frob(Fp2& f, int k) {
  if (k >= 6) return;
  Fp2 g, t;
  fp2_load_const(g, bls::FROB1 + 2 * FPW * k);
  fp2_conj(t, f);
  fp2_mul(f, t, g);
}
DEVN void g6_frob2(Fp2& f, int k) {
  if (k >= 6) return;
  Fp g;
  fp_load_const(g, bls::FROB2 + FPW * k);
  fp2_mul_fp(f, f, g);
}

// Granger-Scott squaring (cyclotomic subgroup).  Pairs (w^p, w^(p+3)) form
// Fp4 elements; lane p < 3 squares pair p, then every lane rebuilds its slot:
//   w0' = 3 s00 - 2 w0,   w3' = 3 s01 + 2 w3   (pair 0)
//   w2' = 3 s10 - 2 w2,   w5' = 3 s11 + 2 w5   (pair 1)
//   w1' = 3 xi s21 + 2 w1, w4' = 3 s20 - 2 w4  (pair 2)
DEVN void g6_cyc_sqr(Fp2& f, int k, uint32_t* lds) {
  uint32_t* A = lds;            // inputs
  uint32_t* S = lds + 6 * G6_F2;   // pair results: s_p0 at 2p, s_p1 at 2p + 1
  __syncthreads();
  if (k < 6) g6_put(A + G6_F2 * k, f);
  __syncthreads();
  if (k < 3) {
    Fp2 a, b, s0, s1;
    g6_get(a, A + G6_F2 * k);
    g6_get(b, A + G6_F2 * (k + 3));
    fp4_sqr(s0, s1, a, b);
    g6_put(S + G6_F2 * (2 * k), s0);
    g6_put(S + G6_F2 * (2 * k + 1), s1);
  }
  __syncthreads();
  if (k < 6) {
    // which pair result feeds slot k, which component, sign, and xi factor
    const int pair = (k == 0 || k == 3) ? 0 : ((k == 2 || k == 5) ? 1 : 2);
    const int comp = (k == 3 || k == 5 || k == 1) ? 1 : 0;
    Fp2 s, t;
    g6_get(s, S + G6_F2 * (2 * pair + comp));
    if (k == 1) fp2_mul_xi(s, s);
    fp2_dbl(t, s);
    fp2_add(t, t, s);  // 3 s
    Fp2 w2;
    fp2_dbl(w2, f);
    if (k == 0 || k == 2 || k == 4)
      fp2_sub(f, t, w2);
    else
      fp2_add(f, t, w2);
  }
}

// Inverse: lane 0 gathers the element, inverts it in the tower, scatters.
DEVN void g6_inv(Fp2& f, int k, uint32_t* lds) {
  uint32_t* A = lds;
  __syncthreads();
  if (k < 6) g6_put(A + G6_F2 * k, f);
  __syncthreads();
  if (k == 0) {
    Fp12 x, y;
    g6_get(x.c0.c0, A + 0);
    g6_get(x.c1.c0, A + G6_F2);
    g6_get(x.c0.c1, A + 2 * G6_F2);
    g6_get(x.c1.c1, A + 3 * G6_F2);
    g6_get(x.c0.c2, A + 4 * G6_F2);
    g6_get(x.c1.c2, A + 5 * G6_F2);
    fp12_inv(y, x);
    g6_put(A + 0, y.c0.c0);
    g6_put(A + G6_F2, y.c1.c0);
    g6_put(A + 2 * G6_F2, y.c0.c1);
    g6_put(A + 3 * G6_F2, y.c1.c1);
    g6_put(A + 4 * G6_F2, y.c0.c2);
    g6_put(A + 5 * G6_F2, y.c1.c2);
  }
  __syncthreads();
  if (k < 6) g6_get(f, A + G6_F2 * k);
}

// a^x (x = -|x|) for a cyclotomic element
DEVN void g6_exp_by_x(Fp2& r, const Fp2& a, int k, uint32_t* lds) {
  Fp2 acc = a;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    g6_cyc_sqr(acc, k, lds);
    if ((BLS_X_ABS >> i) & 1) g6_mul(acc, a, k, lds);
  }
  r = acc;
  g6_conj(r, k);
}

// Same chain as final_exponentiation() in pairing.hpp, slot-parallel.
DEVN void g6_final_exp(Fp2& r, const Fp2& f, int k, uint32_t* lds) {
  Fp2 f1, t, a, b, c;
  t = f;
  g6_inv(t, k, lds);
  f1 = f;
  g6_conj(f1, k);
  g6_mul(f1, t, k, lds);           // f^(p^6 - 1)
  t = f1;
  g6_frob2(t, k);
  g6_mul(f1, t, k, lds);           // ^(p^2 + 1)
  g6_exp_by_x(a, f1, k, lds);
  t = f1;
  g6_conj(t, k);
  g6_mul(a, t, k, lds);            // a = f1^(x-1)
  g6_exp_by_x(t, a, k, lds);
  b = a;
  g6_conj(b, k);
  g6_mul(t, b, k, lds);
  a = t;                           // a = f1^((x-1)^2)
  g6_exp_by_x(b, a, k, lds);
  t = a;
  g6_frob(t, k);
  g6_mul(b, t, k, lds);            // b = a^(x+p)
  g6_exp_by_x(c, b, k, lds);
  g6_exp_by_x(c, c, k, lds);
  t = b;
  g6_frob2(t, k);
  g6_mul(c, t, k, lds);
  t = b;
  g6_conj(t, k);
  g6_mul(c, t, k, lds);            // c = b^(x^2 + p^2 - 1)
  t = f1;
  g6_cyc_sqr(t, k, lds);
  g6_mul(t, f1, k, lds);           // f1^3
  g6_mul(c, t, k, lds);
  r = c;
}

// Product of two Miller loops with prepared lines, conjugated (x < 0).
// Pairs with skip[] set contribute 1.
DEVN void g6_miller2(Fp2& f, int k, uint32_t* lds, const G1A* ps, const bool* skip,
                     const Line* const* lines) {
  fp2_zero(f);
  if (k == 0) fp2_one(f);
  int idx = 0;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    if (i != 62) g6_sqr(f, k, lds);
    const int steps = ((BLS_X_ABS >> i) & 1) ? 2 : 1;
    for (int s = 0; s < steps; s++) {
#pragma unroll 1
      for (int q = 0; q < 2; q++) {
        Fp2 l0, l2;
        Fp l3;
        if (skip[q]) {
          // the constant line 1
          fp2_one(l0);
          fp2_zero(l2);
          fp_zero(l3);
        } else {
          const Line* ln = lines[q] + idx;
          l0 = ln->c;
          fp2_mul_fp(l2, ln->lambda, ps[q].x);
          fp2_neg(l2, l2);
          l3 = ps[q].y;
        }
        g6_mul_line(f, l0, l2, l3, k, lds);
      }
      idx++;
    }
  }
  g6_conj(f, k);
}

// 1 when every lane's slot matches the identity (result valid on all lanes)
DEVN bool g6_is_one(const Fp2& f, int k, uint32_t* lds) {
  uint32_t* F = lds + 2 * 6 * G6_F2;
  __syncthreads();
  if (k < 6) {
    bool ok;
    if (k == 0) {
      Fp one;
      fp_one(one);
      ok = fp_eq(f.c0, one) && fp_is_zero(f.c1);
    } else {
      ok = fp2_is_zero(f);
    }
    F[k] = ok ? 1u : 0u;
  }
  __syncthreads();
  bool all = true;
  for (int i = 0; i < 6; i++) all = all && F[i] != 0;
  return all;
}
