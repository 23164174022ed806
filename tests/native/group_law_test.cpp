// CPU check of the G1 group law's lazy forms (cubbs-cred_amd/csrc/curve.hpp),
// compiled for the host with g++ (UBSan catches any signed 32-bit overflow
// in the limb passes):
//  * g1_dbl_in (lazy sums, one signed limb pass per output) against
//    dbl-2009-l written here with normalized additions only;
//  * g1_madd_in and g1_add_in (lazy sums) against madd-2007-bl and
//    add-2007-bl written with normalized additions, equal and opposite
//    inputs included;
//  * g1z_madd (XYZZ, madd-2008-s, lazily negated y) against g1_madd_in
//    (Jacobian, madd-2007-bl), over running sums;
//  * fp_reduce_multiple on values below 26p.
// Inputs are random values below 2p, values with every limb at or near its
// maximum and near zero, and 0, p, 2p - 1.  Results are compared mod p, and
// outputs must be normalized (limbs below 2^28) and below 2p.  Formulas are
// compared as rational functions, so the inputs need not lie on the curve.
// The multiples of p added in each signed limb pass follow from the bounds
// stated beside the pass (every product and normalized difference is below
// 2p); random inputs rarely come near those bounds, so this test guards the
// formulas and the reductions rather than the size of each offset.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "../../cubbs-cred_amd/csrc/curve.hpp"

static bool eq_mod_p(const Fp& a, const Fp& b) {
  Fp x, y;
  fp_canon(x, a);
  fp_canon(y, b);
  return std::memcmp(&x, &y, sizeof x) == 0;
}
static bool normalized_below_2p(const Fp& a) {
  for (int k = 0; k < 14; k++)
    if (a.v[k] >> 28) return false;
  for (int k = 13; k >= 0; k--)
    if (a.v[k] != bls::P2_28[k]) return a.v[k] < bls::P2_28[k];
  return false;
}
static void value_below_2p(std::mt19937_64& g, Fp& a, int mode) {
  for (;;) {
    for (int k = 0; k < 14; k++) {
      uint32_t v = (uint32_t)g() & FP_MASK;
      if (mode == 1) v = FP_MASK - ((uint32_t)g() & 15);
      if (mode == 2) v = (uint32_t)g() & 15;
      a.v[k] = v;
    }
    a.v[13] = (uint32_t)(g() % (bls::P2_28[13] + 1));
    if (mode == 1) a.v[13] = bls::P2_28[13] - (uint32_t)(g() % 2);
    if (normalized_below_2p(a)) return;
  }
}
static void special(Fp& a, int which) {
  fp_zero(a);
  if (which == 1)
    for (int k = 0; k < 14; k++) a.v[k] = bls::P28[k];
  if (which == 2) {
    for (int k = 0; k < 14; k++) a.v[k] = bls::P2_28[k];
    a.v[0] -= 1;
  }
}
// dbl-2009-l with normalized additions (the reference)
static void dbl_reference(G1J& p) {
  Fp A, B, C, D, E, F, t;
  fp_sqr(A, p.x);
  fp_sqr(B, p.y);
  fp_sqr(C, B);
  fp_add(t, p.x, B);
  fp_sqr(t, t);
  fp_sub(t, t, A);
  fp_sub(t, t, C);
  fp_dbl(D, t);
  fp_dbl(E, A);
  fp_add(E, E, A);
  fp_sqr(F, E);
  Fp x3, y3, z3;
  fp_sub(x3, F, D);
  fp_sub(x3, x3, D);
  fp_sub(t, D, x3);
  fp_mul(y3, E, t);
  fp_dbl(C, C);
  fp_dbl(C, C);
  fp_dbl(C, C);
  fp_sub(y3, y3, C);
  fp_mul(z3, p.y, p.z);
  fp_dbl(z3, z3);
  p.x = x3;
  p.y = y3;
  p.z = z3;
}
// madd-2007-bl and add-2007-bl with normalized additions (the references)
static void madd_reference(G1J& p, const G1A& q) {
  if (fp_is_zero(p.z)) {
    pt_from_affine(p, q);
    return;
  }
  Fp Z1Z1, U2, S2, H, r, HH, I, J, V, t, x3, y3, z3;
  fp_sqr(Z1Z1, p.z);
  fp_mul(U2, q.x, Z1Z1);
  fp_mul(S2, q.y, p.z);
  fp_mul(S2, S2, Z1Z1);
  fp_sub(H, U2, p.x);
  fp_sub(r, S2, p.y);
  if (fp_is_zero(H)) {
    G1J qj;
    pt_from_affine(qj, q);
    if (fp_is_zero(r)) {
      pt_dbl(p, qj);
    } else {
      pt_set_inf(p);
    }
    return;
  }
  fp_dbl(r, r);
  fp_sqr(HH, H);
  fp_add(t, p.z, H);
  fp_sqr(t, t);
  fp_sub(t, t, Z1Z1);
  fp_sub(z3, t, HH);
  fp_dbl(I, HH);
  fp_dbl(I, I);
  fp_mul(J, H, I);
  fp_mul(V, p.x, I);
  fp_sqr(x3, r);
  fp_sub(x3, x3, J);
  fp_sub(x3, x3, V);
  fp_sub(x3, x3, V);
  fp_sub(t, V, x3);
  fp_mul(y3, r, t);
  fp_mul(t, p.y, J);
  fp_dbl(t, t);
  fp_sub(y3, y3, t);
  p.x = x3;
  p.y = y3;
  p.z = z3;
}
static void add_reference(G1J& p, const G1J& q) {
  G1J r;
  pt_add(r, p, q);  // the out-of-line form: add-2007-bl, normalized
  p = r;
}
static bool same_point(const G1J& a, const G1J& b) {
  // projective equality X1 Z2^2 = X2 Z1^2, Y1 Z2^3 = Y2 Z1^3 (or both identity)
  const bool ia = fp_is_zero(a.z), ib = fp_is_zero(b.z);
  if (ia || ib) return ia && ib;
  Fp za2, zb2, za3, zb3, l, r;
  fp_sqr(za2, a.z);
  fp_sqr(zb2, b.z);
  fp_mul(za3, za2, a.z);
  fp_mul(zb3, zb2, b.z);
  fp_mul(l, a.x, zb2);
  fp_mul(r, b.x, za2);
  if (!eq_mod_p(l, r)) return false;
  fp_mul(l, a.y, zb3);
  fp_mul(r, b.y, za3);
  return eq_mod_p(l, r);
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 200000;
  std::mt19937_64 g(20261204);
  long bad_dbl = 0, bad_madd = 0, bad_red = 0;
  for (long it = 0; it < n; it++) {
    G1J p;
    const int mode = (int)(it % 3);
    value_below_2p(g, p.x, mode);
    value_below_2p(g, p.y, (int)((mode + it / 3) % 3));
    value_below_2p(g, p.z, (int)((mode + 2 * (it / 3)) % 3));
    if (it < 27) {
      special(p.x, (int)(it % 3));
      special(p.y, (int)(it / 3 % 3));
      special(p.z, (int)(it / 9 % 3));
    }
    G1J a = p, b = p;
    const int chain = it % 997 == 0 ? 100 : 1;
    for (int r = 0; r < chain; r++) {
      g1_dbl_in(a);
      dbl_reference(b);
      if (!normalized_below_2p(a.x) || !normalized_below_2p(a.y) || !normalized_below_2p(a.z)) bad_dbl++;
    }
    if (!eq_mod_p(a.x, b.x) || !eq_mod_p(a.y, b.y) || !eq_mod_p(a.z, b.z)) bad_dbl++;
  }
  // running sums of affine terms, some negated, in both coordinate systems
  for (long it = 0; it < n / 20; it++) {
    G1Z z;
    g1z_set_inf(z);
    G1J j;
    pt_set_inf(j);
    const int terms = 1 + (int)(g() % 24);
    for (int t = 0; t < terms; t++) {
      G1A q;
      value_below_2p(g, q.x, (int)(g() % 3));
      value_below_2p(g, q.y, (int)(g() % 3));
      const bool neg = (g() & 1) != 0;
      G1A qn = q;
      if (neg) fp_neg(qn.y, qn.y);
      g1_madd_in(j, qn);
      G1A ql = q;
      fp_cond_neg_lazy(ql.y, neg);
      g1z_madd(z, ql);
      if (!normalized_below_2p(z.x) || !normalized_below_2p(z.y) || !normalized_below_2p(z.zz) ||
          !normalized_below_2p(z.zzz))
        bad_madd++;
    }
    G1J c;
    g1z_to_jacobian(c, z);
    if (!same_point(c, j)) bad_madd++;
  }
  // Jacobian mixed and full additions against the references, with equal
  // and opposite inputs now and then (the doubling and identity cases)
  long bad_add = 0;
  for (long it = 0; it < n / 2; it++) {
    G1J p, q;
    const int mode = (int)(it % 3);
    value_below_2p(g, p.x, mode);
    value_below_2p(g, p.y, (int)(g() % 3));
    value_below_2p(g, p.z, (int)(g() % 3));
    value_below_2p(g, q.x, (int)(g() % 3));
    value_below_2p(g, q.y, (int)(g() % 3));
    value_below_2p(g, q.z, (int)(g() % 3));
    if (it < 27) {
      special(p.x, (int)(it % 3));
      special(q.y, (int)(it / 3 % 3));
      special(p.z, (int)(it / 9 % 3));
    }
    G1A qa;
    qa.x = q.x;
    qa.y = q.y;
    if (it % 50 == 7) {  // p equal to the affine q, or to its negative
      p.x = qa.x;
      p.y = qa.y;
      fp_one(p.z);
      if (it % 100 == 7) fp_neg(p.y, p.y);
    }
    G1J a = p, b = p;
    g1_madd_in(a, qa);
    madd_reference(b, qa);
    if (!eq_mod_p(a.x, b.x) || !eq_mod_p(a.y, b.y) || !eq_mod_p(a.z, b.z) || !normalized_below_2p(a.x) ||
        !normalized_below_2p(a.y) || !normalized_below_2p(a.z))
      bad_add++;
    if (it % 50 == 11) q = p;  // equal Jacobian inputs
    G1J c = p, d = p;
    g1_add_in(c, q);
    add_reference(d, q);
    if (!same_point(c, d) || !normalized_below_2p(c.x) || !normalized_below_2p(c.y) ||
        !normalized_below_2p(c.z))
      bad_add++;
  }
  bad_madd += bad_add;
  // fp_reduce_multiple on normalized values below 26p
  for (long it = 0; it < n; it++) {
    Fp a[13], s;
    const int m = 1 + (int)(g() % 13);
    fp_zero(s);
    uint64_t c = 0;
    for (int k = 0; k < m; k++) value_below_2p(g, a[k], (int)(g() % 3));
    for (int i = 0; i < 14; i++) {  // limb-wise sum with carries: below 26p
      for (int k = 0; k < m; k++) c += a[k].v[i];
      s.v[i] = (uint32_t)(c & FP_MASK);
      c >>= 28;
    }
    Fp r = s, ref;
    fp_reduce_multiple(r);
    fp_zero(ref);
    for (int k = 0; k < m; k++) fp_add(ref, ref, a[k]);
    if (!normalized_below_2p(r) || !eq_mod_p(r, ref)) bad_red++;
  }
  std::printf("%ld %ld %ld\n", bad_dbl, bad_madd, bad_red);
  return (bad_dbl || bad_madd || bad_red) ? 1 : 0;
}
