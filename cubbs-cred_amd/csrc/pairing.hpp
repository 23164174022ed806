// Optimal-ate pairing on BLS12-381.
//
// G2 arguments are "prepared": the 68 affine line functions of the Miller
// loop (63 doublings + 5 additions for |x| = 0xd201000000010000) are stored as
// (lambda, c = lambda * x_T - y_T) in Fp2, 192 bytes per line.  Evaluated at
// P = (xp, yp) in G1 (and multiplied by w^3, a factor the final
// exponentiation removes) a line is the sparse Fp12 element
//     c  +  (-lambda xp) v  +  yp (v w)
// i.e. slots c0.c0, c0.c1 and c1.c1.  Issuer public keys and -G2 stay
// resident in this form.
//
// Final exponentiation: easy part (p^6 - 1)(p^2 + 1), hard part via
//   3 (p^4 - p^2 + 1) / r = (x - 1)^2 (x + p) (x^2 + p^2 - 1) + 3
// with Granger-Scott cyclotomic squarings; the result is the pairing cubed,
// which leaves every "== 1" test unchanged (gcd(3, r) = 1).
#pragma once
#include "curve.hpp"

#define MILLER_LINES 68
static constexpr uint64_t BLS_X_ABS = 0xd201000000010000ull;

struct Line { Fp2 lambda, c; };   // 4 x 14 limbs = 224 bytes

// Compute the lines of a G2 point (affine, not the identity).
DEVN void g2_prepare_lines(Line* out, const G2A& q) {
  Fp2 tx = q.x, ty = q.y;
  int n = 0;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    Fp2 lam, t, d, nx, ny;
    fp2_sqr(t, tx);
    fp2_dbl(lam, t);
    fp2_add(lam, lam, t);
    fp2_dbl(d, ty);
    fp2_inv(d, d);
    fp2_mul(lam, lam, d);
    out[n].lambda = lam;
    fp2_mul(t, lam, tx);
    fp2_sub(out[n].c, t, ty);
    n++;
    fp2_sqr(nx, lam);
    fp2_sub(nx, nx, tx);
    fp2_sub(nx, nx, tx);
    fp2_sub(t, tx, nx);
    fp2_mul(ny, lam, t);
    fp2_sub(ny, ny, ty);
    tx = nx;
    ty = ny;
    if ((BLS_X_ABS >> i) & 1) {
      fp2_sub(t, q.y, ty);
      fp2_sub(d, q.x, tx);
      fp2_inv(d, d);
      fp2_mul(lam, t, d);
      out[n].lambda = lam;
      fp2_mul(t, lam, tx);
      fp2_sub(out[n].c, t, ty);
      n++;
      fp2_sqr(nx, lam);
      fp2_sub(nx, nx, tx);
      fp2_sub(nx, nx, q.x);
      fp2_sub(t, tx, nx);
      fp2_mul(ny, lam, t);
      fp2_sub(ny, ny, ty);
      tx = nx;
      ty = ny;
    }
  }
}

// (x0 + x1 v + x2 v^2) * (a + b v)
DEVN void fp6_mul_by_01(Fp6& r, const Fp6& x, const Fp2& a, const Fp2& b) {
  Fp2 t0, t1, s, u, c0, c1, c2;
  fp2_mul(t0, x.c0, a);
  fp2_mul(t1, x.c1, b);
  fp2_add(s, x.c0, x.c1);
  fp2_add(u, a, b);
  fp2_mul(c1, s, u);
  fp2_sub(c1, c1, t0);
  fp2_sub(c1, c1, t1);
  fp2_mul(c2, x.c2, a);
  fp2_add(c2, c2, t1);
  fp2_mul(c0, x.c2, b);
  fp2_mul_xi(c0, c0);
  fp2_add(c0, c0, t0);
  r.c0 = c0;
  r.c1 = c1;
  r.c2 = c2;
}

// (x0 + x1 v + x2 v^2) * (y v), y in Fp
DEVI void fp6_mul_by_v_fp(Fp6& r, const Fp6& x, const Fp& y) {
  Fp2 t;
  fp2_mul_fp(t, x.c2, y);
  fp2_mul_xi(t, t);
  Fp2 n1, n2;
  fp2_mul_fp(n1, x.c0, y);
  fp2_mul_fp(n2, x.c1, y);
  r.c0 = t;
  r.c1 = n1;
  r.c2 = n2;
}

// f *= (l0 + l1 w) with l0 = (c, m, 0), l1 = (0, yp, 0), m = -lambda xp.
DEVN void fp12_mul_by_line(Fp12& f, const Fp2& c, const Fp2& m, const Fp& yp) {
  Fp6 t0, t1, s;
  fp6_mul_by_01(t0, f.c0, c, m);
  fp6_mul_by_v_fp(t1, f.c1, yp);
  // (f0 + f1)(l0 + l1) with l0 + l1 = (c, m + yp, 0)
  Fp2 m2 = m;
  fp_add(m2.c0, m2.c0, yp);
  fp6_add(s, f.c0, f.c1);
  fp6_mul_by_01(s, s, c, m2);
  fp6_sub(s, s, t0);
  fp6_sub(f.c1, s, t1);
  fp6_mul_v(t1, t1);
  fp6_add(f.c0, t0, t1);
}

DEVN void line_eval_mul(Fp12& f, const Line* ln, const G1A& p) {
  Fp2 m;
  fp2_mul_fp(m, ln->lambda, p.x);
  fp2_neg(m, m);
  fp12_mul_by_line(f, ln->c, m, p.y);
}

// f = prod_k Miller_{|x|}(P_k, Q_k), conjugated because x < 0.  Pairs whose
// G1 point is the identity are skipped (their value is 1).
template <int K>
DEVI void miller_loop_prepared(Fp12& f, const G1A* ps, const bool* skip, const Line* const* lines) {
  fp12_one(f);
  int idx = 0;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    if (i != 62) fp12_sqr(f, f);
#pragma unroll
    for (int k = 0; k < K; k++)
      if (!skip[k]) line_eval_mul(f, lines[k] + idx, ps[k]);
    idx++;
    if ((BLS_X_ABS >> i) & 1) {
#pragma unroll
      for (int k = 0; k < K; k++)
        if (!skip[k]) line_eval_mul(f, lines[k] + idx, ps[k]);
      idx++;
    }
  }
  fp12_conj(f, f);
}

// ------------------------------------------------------------------ final exponentiation
DEVN void fp4_sqr(Fp2& c0, Fp2& c1, const Fp2& a, const Fp2& b) {
  Fp2 t0, t1, t2;
  fp2_sqr(t0, a);
  fp2_sqr(t1, b);
  fp2_mul_xi(t2, t1);
  fp2_add(c0, t2, t0);
  fp2_add(t2, a, b);
  fp2_sqr(t2, t2);
  fp2_sub(t2, t2, t0);
  fp2_sub(c1, t2, t1);
}

// Granger-Scott squaring, valid in the cyclotomic subgroup.
DEVN void fp12_cyclotomic_sqr(Fp12& r, const Fp12& f) {
  Fp2 z0 = f.c0.c0, z4 = f.c0.c1, z3 = f.c0.c2, z2 = f.c1.c0, z1 = f.c1.c1, z5 = f.c1.c2;
  Fp2 t0, t1, t2, t3;
  fp4_sqr(t0, t1, z0, z1);
  fp2_sub(z0, t0, z0);
  fp2_dbl(z0, z0);
  fp2_add(z0, z0, t0);
  fp2_add(z1, t1, z1);
  fp2_dbl(z1, z1);
  fp2_add(z1, z1, t1);
  fp4_sqr(t0, t1, z2, z3);
  fp4_sqr(t2, t3, z4, z5);
  fp2_sub(z4, t0, z4);
  fp2_dbl(z4, z4);
  fp2_add(z4, z4, t0);
  fp2_add(z5, t1, z5);
  fp2_dbl(z5, z5);
  fp2_add(z5, z5, t1);
  fp2_mul_xi(t0, t3);
  fp2_add(z2, t0, z2);
  fp2_dbl(z2, z2);
  fp2_add(z2, z2, t0);
  fp2_sub(z3, t2, z3);
  fp2_dbl(z3, z3);
  fp2_add(z3, z3, t2);
  r.c0.c0 = z0;
  r.c0.c1 = z4;
  r.c0.c2 = z3;
  r.c1.c0 = z2;
  r.c1.c1 = z1;
  r.c1.c2 = z5;
}

// a^x for a in the cyclotomic subgroup, x = -|x|
DEVN void fp12_exp_by_x(Fp12& r, const Fp12& a) {
  Fp12 acc = a;
#pragma unroll 1
  for (int i = 62; i >= 0; i--) {
    fp12_cyclotomic_sqr(acc, acc);
    if ((BLS_X_ABS >> i) & 1) fp12_mul(acc, acc, a);
  }
  fp12_conj(r, acc);
}

DEVN void final_exponentiation(Fp12& r, const Fp12& f) {
  Fp12 f1, t, a, b, c;
  // easy part
  fp12_inv(t, f);
  fp12_conj(f1, f);
  fp12_mul(f1, f1, t);
  fp12_frob2(t, f1);
  fp12_mul(f1, t, f1);
  // a = f1^((x-1)^2)
  fp12_exp_by_x(a, f1);
  fp12_conj(t, f1);
  fp12_mul(a, a, t);
  fp12_exp_by_x(t, a);
  fp12_conj(b, a);
  fp12_mul(a, t, b);
  // b = a^(x+p)
  fp12_exp_by_x(b, a);
  fp12_frob(t, a);
  fp12_mul(b, b, t);
  // c = b^(x^2 + p^2 - 1)
  fp12_exp_by_x(c, b);
  fp12_exp_by_x(c, c);
  fp12_frob2(t, b);
  fp12_mul(c, c, t);
  fp12_conj(t, b);
  fp12_mul(c, c, t);
  // * f1^3
  fp12_cyclotomic_sqr(t, f1);
  fp12_mul(t, t, f1);
  fp12_mul(r, c, t);
}
