// RFC 9380 hash_to_curve for G1 (suite BLS12381G1_XMD:SHA-256_SSWU_RO_) on the
// device, used to create the BBS generators (crates/cubbs-bbs-protocol/src/
// bbsplus/generators.rs:119-148): hash_to_field (expand_message_xmd, 128
// bytes, two elements mod p), simplified SWU onto the 11-isogenous curve E',
// the 11-isogeny onto E, point addition and cofactor clearing by
// h_eff = 1 - x = 0xd201000000010001.
#pragma once
#include "curve.hpp"
#include "sha.hpp"

// expand_message_xmd(msg = 48 bytes given as 12 big-endian words, dst, 128)
DEVN void xmd128_from_words(uint32_t out[32], const uint32_t* msg_words, uint32_t msg_words_n,
                            const uint8_t* dst, uint32_t dst_len) {
  Sha256 s;
  sha_init_after_zpad(s);
  for (uint32_t k = 0; k < msg_words_n; k++) sha_word(s, msg_words[k]);
  xmd_b0_suffix(s, 128, dst, dst_len);
  uint32_t b0[8], bi[8], x[8];
#pragma unroll
  for (int k = 0; k < 8; k++) b0[k] = s.h[k];
  xmd_bi(bi, b0, 1, dst, dst_len);
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = bi[k];
  for (uint32_t i = 2; i <= 4; i++) {
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = b0[k] ^ bi[k];
    xmd_bi(bi, x, i, dst, dst_len);
#pragma unroll
    for (int k = 0; k < 8; k++) out[8 * (i - 1) + k] = bi[k];
  }
}

DEVN void fp_horner_const(Fp& r, const uint32_t* coeffs, int n, const Fp& x) {
  Fp acc, c;
  fp_zero(acc);
#pragma unroll 1
  for (int i = n - 1; i >= 0; i--) {
    fp_mul(acc, acc, x);
    fp_load_const(c, coeffs + FPW * i);
    fp_add(acc, acc, c);
  }
  r = acc;
}

// map_to_curve: simplified SWU on E' then the 11-isogeny; result affine on E.
DEVN void map_to_curve_g1(G1A& out, const Fp& u) {
  Fp z, A, B, tv1, u2, t, x1, gx1, y, x;
  fp_load_const(z, bls::SSWU_Z);
  fp_load_const(A, bls::SSWU_A);
  fp_load_const(B, bls::SSWU_B);
  fp_sqr(u2, u);
  fp_mul(tv1, u2, z);
  fp_sqr(t, tv1);
  fp_add(tv1, t, tv1);          // Z^2 u^4 + Z u^2
  if (fp_is_zero(tv1)) {
    fp_load_const(x1, bls::SSWU_B_OVER_ZA);
  } else {
    Fp one, mba;
    fp_inv(tv1, tv1);
    fp_one(one);
    fp_add(tv1, tv1, one);
    fp_load_const(mba, bls::SSWU_MINUS_B_OVER_A);
    fp_mul(x1, mba, tv1);
  }
  fp_sqr(gx1, x1);
  fp_add(gx1, gx1, A);
  fp_mul(gx1, gx1, x1);
  fp_add(gx1, gx1, B);
  if (fp_sqrt(y, gx1)) {
    x = x1;
  } else {
    Fp gx2;
    fp_mul(x, z, u2);
    fp_mul(x, x, x1);
    fp_sqr(gx2, x);
    fp_add(gx2, gx2, A);
    fp_mul(gx2, gx2, x);
    fp_add(gx2, gx2, B);
    fp_sqrt(y, gx2);
  }
  if (fp_sgn0(u) != fp_sgn0(y)) fp_neg(y, y);
  Fp xn, xd, yn, yd;
  fp_horner_const(xn, bls::ISO_XNUM, bls::ISO_XNUM_LEN, x);
  fp_horner_const(xd, bls::ISO_XDEN, bls::ISO_XDEN_LEN, x);
  fp_horner_const(yn, bls::ISO_YNUM, bls::ISO_YNUM_LEN, x);
  fp_horner_const(yd, bls::ISO_YDEN, bls::ISO_YDEN_LEN, x);
  // one inversion for both denominators
  Fp dd, inv;
  fp_mul(dd, xd, yd);
  fp_inv(inv, dd);
  Fp ixd, iyd;
  fp_mul(ixd, inv, yd);
  fp_mul(iyd, inv, xd);
  fp_mul(out.x, xn, ixd);
  fp_mul(t, yn, iyd);
  fp_mul(out.y, y, t);
}

// hash_to_curve(msg, dst) with the message given as 32-bit big-endian words.
DEVN void hash_to_curve_g1_words(G1J& out, const uint32_t* msg_words, uint32_t n,
                                 const uint8_t* dst, uint32_t dst_len) {
  uint32_t ub[32];
  xmd128_from_words(ub, msg_words, n, dst, dst_len);
  G1A q0, q1;
#pragma unroll
  for (int e = 0; e < 2; e++) {
    uint32_t v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = ub[16 * e + 15 - k];  // big-endian words -> LE limbs
    Fp u;
    fp_reduce_wide_to_mont(u, v);
    if (e == 0)
      map_to_curve_g1(q0, u);
    else
      map_to_curve_g1(q1, u);
  }
  G1J s;
  pt_from_affine(s, q0);
  pt_madd(s, s, q1);
  pt_mul_u64(out, s, 0xd201000000010001ull);
}
