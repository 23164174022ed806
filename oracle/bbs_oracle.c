/*
 * CPU oracle for the BBS (BLS12-381, SHA-256) hot path -- TEST INFRASTRUCTURE.
 *
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg use
 * this file; the HIP library is an independent implementation (32-bit limbs,
 * device code) checked against it.
 *
 * Layout of this file:
 *   1. SHA-256 (FIPS 180-4), expand_message_xmd (RFC 9380 5.3.1), ChaCha20
 *   2. Montgomery arithmetic for Fp (6x64) and Fr (4x64)
 *   3. Fp2 / Fp6 / Fp12 tower, Frobenius constants derived at init
 *   4. G1 / G2 Jacobian arithmetic, ZCash compression codec, subgroup checks
 *   5. Optimal-ate pairing (affine Miller loop, x-chain final exponentiation)
 *   6. RFC 9380 hash_to_curve for G1 (SSWU on the 11-isogenous curve)
 *   7. BBS: generators, domain, Sign, Verify, ProofGen, ProofVerify
 * Reference citations are given on each BBS function.
 */
#include "bbs_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "h2c_iso_consts.h"

typedef unsigned __int128 u128;

/* ===================================================================== 1. hashing */
static const uint32_t SHA_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

typedef struct {
  uint32_t h[8];
  uint8_t buf[64];
  size_t buflen;
  uint64_t total;
} sha256_ctx;

#define ROR32(x, n) (((x) >> (n)) | ((x) << (32 - (n))))

static void sha256_compress(uint32_t h[8], const uint8_t blk[64]) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = ((uint32_t)blk[4 * i] << 24) | ((uint32_t)blk[4 * i + 1] << 16) |
           ((uint32_t)blk[4 * i + 2] << 8) | blk[4 * i + 3];
  for (int i = 16; i < 64; i++) {
    uint32_t s0 = ROR32(w[i - 15], 7) ^ ROR32(w[i - 15], 18) ^ (w[i - 15] >> 3);
    uint32_t s1 = ROR32(w[i - 2], 17) ^ ROR32(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
  for (int i = 0; i < 64; i++) {
    uint32_t S1 = ROR32(e, 6) ^ ROR32(e, 11) ^ ROR32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = hh + S1 + ch + SHA_K[i] + w[i];
    uint32_t S0 = ROR32(a, 2) ^ ROR32(a, 13) ^ ROR32(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + mj;
    hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

static void sha256_init(sha256_ctx *c) {
  static const uint32_t iv[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  memcpy(c->h, iv, sizeof iv);
  c->buflen = 0;
  c->total = 0;
}

static void sha256_update(sha256_ctx *c, const uint8_t *p, size_t n) {
  c->total += n;
  while (n) {
    size_t take = 64 - c->buflen;
    if (take > n) take = n;
    memcpy(c->buf + c->buflen, p, take);
    c->buflen += take;
    p += take;
    n -= take;
    if (c->buflen == 64) {
      sha256_compress(c->h, c->buf);
      c->buflen = 0;
    }
  }
}

static void sha256_final(sha256_ctx *c, uint8_t out[32]) {
  uint64_t bits = c->total * 8;
  uint8_t pad = 0x80;
  sha256_update(c, &pad, 1);
  uint8_t z = 0;
  while (c->buflen != 56) sha256_update(c, &z, 1);
  uint8_t lenb[8];
  for (int i = 0; i < 8; i++) lenb[i] = (uint8_t)(bits >> (56 - 8 * i));
  sha256_update(c, lenb, 8);
  for (int i = 0; i < 8; i++) {
    out[4 * i] = (uint8_t)(c->h[i] >> 24);
    out[4 * i + 1] = (uint8_t)(c->h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(c->h[i] >> 8);
    out[4 * i + 3] = (uint8_t)c->h[i];
  }
}

void orc_sha256(const uint8_t *msg, size_t len, uint8_t out[32]) {
  sha256_ctx c;
  sha256_init(&c);
  sha256_update(&c, msg, len);
  sha256_final(&c, out);
}

/* RFC 9380 5.3.1 expand_message_xmd with SHA-256 (b_in_bytes 32, s_in_bytes 64). */
int orc_expand_message_xmd(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                           uint8_t *out, size_t out_len) {
  if (dst_len > 255 || out_len > 255 * 32 || out_len > 65535) return ORC_ERR_INPUT;
  size_t ell = (out_len + 31) / 32;
  uint8_t zpad[64] = {0};
  uint8_t lib[2] = {(uint8_t)(out_len >> 8), (uint8_t)out_len};
  uint8_t zero = 0, dl = (uint8_t)dst_len;
  uint8_t b0[32], bi[32];
  sha256_ctx c;
  sha256_init(&c);
  sha256_update(&c, zpad, 64);
  sha256_update(&c, msg, len);
  sha256_update(&c, lib, 2);
  sha256_update(&c, &zero, 1);
  sha256_update(&c, dst, dst_len);
  sha256_update(&c, &dl, 1);
  sha256_final(&c, b0);
  uint8_t one = 1;
  sha256_init(&c);
  sha256_update(&c, b0, 32);
  sha256_update(&c, &one, 1);
  sha256_update(&c, dst, dst_len);
  sha256_update(&c, &dl, 1);
  sha256_final(&c, bi);
  size_t off = 0;
  for (size_t i = 1; i <= ell; i++) {
    size_t take = out_len - off < 32 ? out_len - off : 32;
    memcpy(out + off, bi, take);
    off += take;
    if (i == ell) break;
    uint8_t x[32];
    for (int k = 0; k < 32; k++) x[k] = b0[k] ^ bi[k];
    uint8_t idx = (uint8_t)(i + 1);
    sha256_init(&c);
    sha256_update(&c, x, 32);
    sha256_update(&c, &idx, 1);
    sha256_update(&c, dst, dst_len);
    sha256_update(&c, &dl, 1);
    sha256_final(&c, bi);
  }
  return ORC_OK;
}

/* RFC 8439 ChaCha20 block function. */
#define ROL32(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define QR(a, b, c, d) \
  a += b; d ^= a; d = ROL32(d, 16); c += d; b ^= c; b = ROL32(b, 12); \
  a += b; d ^= a; d = ROL32(d, 8); c += d; b ^= c; b = ROL32(b, 7);
static uint32_t le32(const uint8_t *p) {
  return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
void orc_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                        uint8_t out[64]) {
  uint32_t s[16], x[16];
  s[0] = 0x61707865; s[1] = 0x3320646e; s[2] = 0x79622d32; s[3] = 0x6b206574;
  for (int i = 0; i < 8; i++) s[4 + i] = le32(key + 4 * i);
  s[12] = counter;
  for (int i = 0; i < 3; i++) s[13 + i] = le32(nonce + 4 * i);
  memcpy(x, s, sizeof s);
  for (int i = 0; i < 10; i++) {
    QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
    QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
    QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
    QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
  }
  for (int i = 0; i < 16; i++) {
    uint32_t v = x[i] + s[i];
    out[4 * i] = (uint8_t)v; out[4 * i + 1] = (uint8_t)(v >> 8);
    out[4 * i + 2] = (uint8_t)(v >> 16); out[4 * i + 3] = (uint8_t)(v >> 24);
  }
}

/* ===================================================================== 2. Montgomery */
#define NP 6
#define NR 4
typedef struct { uint64_t l[NP]; } fp;
typedef struct { uint64_t l[NR]; } fr;

static const uint64_t P_L[NP] = {0xb9feffffffffaaabULL, 0x1eabfffeb153ffffULL,
                                 0x6730d2a0f6b0f624ULL, 0x64774b84f38512bfULL,
                                 0x4b1ba7b6434bacd7ULL, 0x1a0111ea397fe69aULL};
static const uint64_t R_L[NR] = {0xffffffff00000001ULL, 0x53bda402fffe5bfeULL,
                                 0x3339d80809a1d805ULL, 0x73eda753299d7d48ULL};
static uint64_t P_INV, R_INV;  /* -m^-1 mod 2^64 */
static fp FP_R2, FP_ONE;       /* 2^768 mod p ; Montgomery one */
static fr FR_R2, FR_ONE;
static uint64_t P_MINUS_2[NP], P_PLUS1_DIV4[NP], P_MINUS1_DIV2[NP], P_MINUS3_DIV4[NP];
static uint64_t R_MINUS_2[NR];

static inline int mp_geq(const uint64_t *a, const uint64_t *b, int n) {
  for (int i = n - 1; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] > b[i];
  }
  return 1;
}
static inline uint64_t mp_sub(uint64_t *r, const uint64_t *a, const uint64_t *b, int n) {
  uint64_t borrow = 0;
  for (int i = 0; i < n; i++) {
    u128 d = (u128)a[i] - b[i] - borrow;
    r[i] = (uint64_t)d;
    borrow = (uint64_t)(d >> 64) & 1;
  }
  return borrow;
}
static inline uint64_t mp_add(uint64_t *r, const uint64_t *a, const uint64_t *b, int n) {
  u128 c = 0;
  for (int i = 0; i < n; i++) {
    c += (u128)a[i] + b[i];
    r[i] = (uint64_t)c;
    c >>= 64;
  }
  return (uint64_t)c;
}

/* CIOS Montgomery multiplication: r = a*b*2^(-64n) mod m. */
static inline void mont_mul(uint64_t *r, const uint64_t *a, const uint64_t *b, const uint64_t *m,
                            uint64_t minv, int n) {
  uint64_t t[NP + 2];
  memset(t, 0, sizeof t);
  for (int i = 0; i < n; i++) {
    u128 c = 0;
    for (int j = 0; j < n; j++) {
      c += (u128)a[j] * b[i] + t[j];
      t[j] = (uint64_t)c;
      c >>= 64;
    }
    c += t[n];
    t[n] = (uint64_t)c;
    t[n + 1] = (uint64_t)(c >> 64);
    uint64_t q = t[0] * minv;
    c = (u128)q * m[0] + t[0];
    c >>= 64;
    for (int j = 1; j < n; j++) {
      c += (u128)q * m[j] + t[j];
      t[j - 1] = (uint64_t)c;
      c >>= 64;
    }
    c += t[n];
    t[n - 1] = (uint64_t)c;
    t[n] = t[n + 1] + (uint64_t)(c >> 64);
  }
  if (t[n] || mp_geq(t, m, n))
    mp_sub(r, t, m, n);
  else
    memcpy(r, t, 8 * n);
}

static void fp_mul(fp *r, const fp *a, const fp *b) { mont_mul(r->l, a->l, b->l, P_L, P_INV, NP); }
static void fp_sqr(fp *r, const fp *a) { mont_mul(r->l, a->l, a->l, P_L, P_INV, NP); }
static void fp_add(fp *r, const fp *a, const fp *b) {
  uint64_t c = mp_add(r->l, a->l, b->l, NP);
  if (c || mp_geq(r->l, P_L, NP)) mp_sub(r->l, r->l, P_L, NP);
}
static void fp_sub(fp *r, const fp *a, const fp *b) {
  if (mp_sub(r->l, a->l, b->l, NP)) mp_add(r->l, r->l, P_L, NP);
}
static int fp_is_zero(const fp *a) {
  uint64_t o = 0;
  for (int i = 0; i < NP; i++) o |= a->l[i];
  return o == 0;
}
static void fp_neg(fp *r, const fp *a) {
  if (fp_is_zero(a)) { *r = *a; return; }
  mp_sub(r->l, P_L, a->l, NP);
}
static int fp_eq(const fp *a, const fp *b) { return memcmp(a, b, sizeof(fp)) == 0; }
static void fp_pow(fp *r, const fp *a, const uint64_t *e, int n) {
  fp acc = FP_ONE, base = *a;
  for (int i = n * 64 - 1; i >= 0; i--) {
    fp_sqr(&acc, &acc);
    if ((e[i / 64] >> (i % 64)) & 1) fp_mul(&acc, &acc, &base);
  }
  *r = acc;
}
static void fp_inv(fp *r, const fp *a) { fp_pow(r, a, P_MINUS_2, NP); }
static void fp_from_limbs(fp *r, const uint64_t *l) {  /* canonical -> Montgomery */
  fp t;
  memcpy(t.l, l, sizeof t.l);
  fp_mul(r, &t, &FP_R2);
}
static void fp_to_limbs(uint64_t *l, const fp *a) {
  fp one = {{1, 0, 0, 0, 0, 0}}, t;
  fp_mul(&t, a, &one);
  memcpy(l, t.l, sizeof t.l);
}
static void fp_from_u64(fp *r, uint64_t v) {
  uint64_t l[NP] = {v, 0, 0, 0, 0, 0};
  fp_from_limbs(r, l);
}
static void be_to_limbs(uint64_t *l, const uint8_t *be, int nbytes, int nlimbs) {
  memset(l, 0, 8 * nlimbs);
  for (int i = 0; i < nbytes; i++) {
    int bit = 8 * (nbytes - 1 - i);
    l[bit / 64] |= (uint64_t)be[i] << (bit % 64);
  }
}
static void limbs_to_be(uint8_t *be, const uint64_t *l, int nbytes) {
  for (int i = 0; i < nbytes; i++) {
    int bit = 8 * (nbytes - 1 - i);
    be[i] = (uint8_t)(l[bit / 64] >> (bit % 64));
  }
}
/* returns 0 if the 48-byte big-endian value is >= p */
static int fp_from_be(fp *r, const uint8_t be[48]) {
  uint64_t l[NP];
  be_to_limbs(l, be, 48, NP);
  if (mp_geq(l, P_L, NP)) return 0;
  fp_from_limbs(r, l);
  return 1;
}
static void fp_to_be(uint8_t be[48], const fp *a) {
  uint64_t l[NP];
  fp_to_limbs(l, a);
  limbs_to_be(be, l, 48);
}
static int fp_sqrt(fp *r, const fp *a) {
  fp s, t;
  fp_pow(&s, a, P_PLUS1_DIV4, NP);
  fp_sqr(&t, &s);
  if (!fp_eq(&t, a)) return 0;
  *r = s;
  return 1;
}
/* lexicographically largest: canonical value > (p-1)/2 */
static int fp_lex_largest(const fp *a) {
  uint64_t l[NP];
  fp_to_limbs(l, a);
  if (mp_geq(P_MINUS1_DIV2, l, NP)) return 0;
  return 1;
}
static int fp_sgn0(const fp *a) {
  uint64_t l[NP];
  fp_to_limbs(l, a);
  return (int)(l[0] & 1);
}
static void fp_from_hex(fp *r, const char *hex) {
  uint8_t be[48];
  for (int i = 0; i < 48; i++) {
    unsigned v;
    char tmp[3] = {hex[2 * i], hex[2 * i + 1], 0};
    v = (unsigned)strtoul(tmp, NULL, 16);
    be[i] = (uint8_t)v;
  }
  fp_from_be(r, be);
}
/* Horner reduction of a big-endian byte string of any length mod p. */
static void fp_from_be_wide(fp *r, const uint8_t *be, size_t n) {
  fp acc = {{0}}, b256, d;
  fp_from_u64(&b256, 256);
  for (size_t i = 0; i < n; i++) {
    fp_mul(&acc, &acc, &b256);
    fp_from_u64(&d, be[i]);
    fp_add(&acc, &acc, &d);
  }
  *r = acc;
}

static void fr_mul(fr *r, const fr *a, const fr *b) { mont_mul(r->l, a->l, b->l, R_L, R_INV, NR); }
static void fr_add(fr *r, const fr *a, const fr *b) {
  uint64_t c = mp_add(r->l, a->l, b->l, NR);
  if (c || mp_geq(r->l, R_L, NR)) mp_sub(r->l, r->l, R_L, NR);
}
static void fr_sub(fr *r, const fr *a, const fr *b) {
  if (mp_sub(r->l, a->l, b->l, NR)) mp_add(r->l, r->l, R_L, NR);
}
static int fr_is_zero(const fr *a) { return (a->l[0] | a->l[1] | a->l[2] | a->l[3]) == 0; }
static int fr_eq(const fr *a, const fr *b) { return memcmp(a, b, sizeof(fr)) == 0; }
static void fr_neg(fr *r, const fr *a) {
  if (fr_is_zero(a)) { *r = *a; return; }
  mp_sub(r->l, R_L, a->l, NR);
}
static void fr_inv(fr *r, const fr *a) {
  fr acc = FR_ONE, base = *a;
  for (int i = 255; i >= 0; i--) {
    fr_mul(&acc, &acc, &acc);
    if ((R_MINUS_2[i / 64] >> (i % 64)) & 1) fr_mul(&acc, &acc, &base);
  }
  *r = acc;
}
static void fr_from_limbs(fr *r, const uint64_t *l) {
  fr t;
  memcpy(t.l, l, sizeof t.l);
  fr_mul(r, &t, &FR_R2);
}
static void fr_to_limbs(uint64_t *l, const fr *a) {
  fr one = {{1, 0, 0, 0}}, t;
  fr_mul(&t, a, &one);
  memcpy(l, t.l, sizeof t.l);
}
static void fr_from_u64(fr *r, uint64_t v) {
  uint64_t l[NR] = {v, 0, 0, 0};
  fr_from_limbs(r, l);
}
/* canonical big-endian decode; 0 when value >= r (bls12_381_plus Scalar::from_be_bytes) */
static int fr_from_be(fr *r, const uint8_t be[32]) {
  uint64_t l[NR];
  be_to_limbs(l, be, 32, NR);
  if (mp_geq(l, R_L, NR)) return 0;
  fr_from_limbs(r, l);
  return 1;
}
static void fr_to_be(uint8_t be[32], const fr *a) {
  uint64_t l[NR];
  fr_to_limbs(l, a);
  limbs_to_be(be, l, 32);
}
/* OS2IP(bytes) mod r for any length (Scalar::from_okm for 48 bytes). */
static void fr_from_be_wide(fr *r, const uint8_t *be, size_t n) {
  fr acc = {{0}}, b256, d;
  fr_from_u64(&b256, 256);
  for (size_t i = 0; i < n; i++) {
    fr_mul(&acc, &acc, &b256);
    fr_from_u64(&d, be[i]);
    fr_add(&acc, &acc, &d);
  }
  *r = acc;
}

/* ===================================================================== 3. tower */
typedef struct { fp c0, c1; } fp2;
typedef struct { fp2 c0, c1, c2; } fp6;
typedef struct { fp6 c0, c1; } fp12;

static void fp2_add(fp2 *r, const fp2 *a, const fp2 *b) { fp_add(&r->c0, &a->c0, &b->c0); fp_add(&r->c1, &a->c1, &b->c1); }
static void fp2_sub(fp2 *r, const fp2 *a, const fp2 *b) { fp_sub(&r->c0, &a->c0, &b->c0); fp_sub(&r->c1, &a->c1, &b->c1); }
static void fp2_neg(fp2 *r, const fp2 *a) { fp_neg(&r->c0, &a->c0); fp_neg(&r->c1, &a->c1); }
static void fp2_conj(fp2 *r, const fp2 *a) { r->c0 = a->c0; fp_neg(&r->c1, &a->c1); }
static int fp2_is_zero(const fp2 *a) { return fp_is_zero(&a->c0) && fp_is_zero(&a->c1); }
static int fp2_eq(const fp2 *a, const fp2 *b) { return fp_eq(&a->c0, &b->c0) && fp_eq(&a->c1, &b->c1); }
static void fp2_mul(fp2 *r, const fp2 *a, const fp2 *b) {
  fp v0, v1, s0, s1, t;
  fp_mul(&v0, &a->c0, &b->c0);
  fp_mul(&v1, &a->c1, &b->c1);
  fp_add(&s0, &a->c0, &a->c1);
  fp_add(&s1, &b->c0, &b->c1);
  fp_mul(&t, &s0, &s1);
  fp_sub(&t, &t, &v0);
  fp_sub(&r->c1, &t, &v1);
  fp_sub(&r->c0, &v0, &v1);
}
static void fp2_sqr(fp2 *r, const fp2 *a) {
  fp s, d, m;
  fp_add(&s, &a->c0, &a->c1);
  fp_sub(&d, &a->c0, &a->c1);
  fp_mul(&m, &a->c0, &a->c1);
  fp_mul(&r->c0, &s, &d);
  fp_add(&r->c1, &m, &m);
}
static void fp2_mul_fp(fp2 *r, const fp2 *a, const fp *b) { fp_mul(&r->c0, &a->c0, b); fp_mul(&r->c1, &a->c1, b); }
static void fp2_mul_xi(fp2 *r, const fp2 *a) {  /* * (1 + u) */
  fp t0, t1;
  fp_sub(&t0, &a->c0, &a->c1);
  fp_add(&t1, &a->c0, &a->c1);
  r->c0 = t0;
  r->c1 = t1;
}
static void fp2_inv(fp2 *r, const fp2 *a) {
  fp n, t;
  fp_sqr(&n, &a->c0);
  fp_sqr(&t, &a->c1);
  fp_add(&n, &n, &t);
  fp_inv(&n, &n);
  fp_mul(&r->c0, &a->c0, &n);
  fp_mul(&t, &a->c1, &n);
  fp_neg(&r->c1, &t);
}
static void fp2_pow(fp2 *r, const fp2 *a, const uint64_t *e, int n) {
  fp2 acc, base = *a;
  acc.c0 = FP_ONE;
  memset(&acc.c1, 0, sizeof(fp));
  for (int i = n * 64 - 1; i >= 0; i--) {
    fp2_sqr(&acc, &acc);
    if ((e[i / 64] >> (i % 64)) & 1) fp2_mul(&acc, &acc, &base);
  }
  *r = acc;
}
static void fp2_one(fp2 *r) { r->c0 = FP_ONE; memset(&r->c1, 0, sizeof(fp)); }
/* Square root in Fp2 for p = 3 mod 4 (Adj & Rodriguez-Henriquez, Alg. 9). */
static int fp2_sqrt(fp2 *r, const fp2 *a) {
  if (fp2_is_zero(a)) { *r = *a; return 1; }
  fp2 a1, alpha, x0, t, one, minus_one;
  fp2_one(&one);
  fp2_neg(&minus_one, &one);
  fp2_pow(&a1, a, P_MINUS3_DIV4, NP);
  fp2_mul(&alpha, &a1, a);
  fp2_mul(&alpha, &alpha, &a1);
  fp2_mul(&x0, &a1, a);
  if (fp2_eq(&alpha, &minus_one)) {
    /* x = u * x0 */
    fp2 u = {{{0}}, FP_ONE};
    fp2_mul(&t, &u, &x0);
  } else {
    fp2 b;
    fp2_add(&b, &alpha, &one);
    fp2_pow(&b, &b, P_MINUS1_DIV2, NP);
    fp2_mul(&t, &b, &x0);
  }
  fp2 chk;
  fp2_sqr(&chk, &t);
  if (!fp2_eq(&chk, a)) return 0;
  *r = t;
  return 1;
}
static int fp2_lex_largest(const fp2 *a) {
  if (!fp_is_zero(&a->c1)) return fp_lex_largest(&a->c1);
  return fp_lex_largest(&a->c0);
}

static void fp6_add(fp6 *r, const fp6 *a, const fp6 *b) { fp2_add(&r->c0, &a->c0, &b->c0); fp2_add(&r->c1, &a->c1, &b->c1); fp2_add(&r->c2, &a->c2, &b->c2); }
static void fp6_sub(fp6 *r, const fp6 *a, const fp6 *b) { fp2_sub(&r->c0, &a->c0, &b->c0); fp2_sub(&r->c1, &a->c1, &b->c1); fp2_sub(&r->c2, &a->c2, &b->c2); }
static void fp6_neg(fp6 *r, const fp6 *a) { fp2_neg(&r->c0, &a->c0); fp2_neg(&r->c1, &a->c1); fp2_neg(&r->c2, &a->c2); }
static void fp6_mul(fp6 *r, const fp6 *a, const fp6 *b) {
  fp2 t0, t1, t2, s0, s1, c0, c1, c2;
  fp2_mul(&t0, &a->c0, &b->c0);
  fp2_mul(&t1, &a->c1, &b->c1);
  fp2_mul(&t2, &a->c2, &b->c2);
  fp2_add(&s0, &a->c1, &a->c2); fp2_add(&s1, &b->c1, &b->c2); fp2_mul(&c0, &s0, &s1);
  fp2_sub(&c0, &c0, &t1); fp2_sub(&c0, &c0, &t2); fp2_mul_xi(&c0, &c0); fp2_add(&c0, &c0, &t0);
  fp2_add(&s0, &a->c0, &a->c1); fp2_add(&s1, &b->c0, &b->c1); fp2_mul(&c1, &s0, &s1);
  fp2_sub(&c1, &c1, &t0); fp2_sub(&c1, &c1, &t1); fp2_mul_xi(&s0, &t2); fp2_add(&c1, &c1, &s0);
  fp2_add(&s0, &a->c0, &a->c2); fp2_add(&s1, &b->c0, &b->c2); fp2_mul(&c2, &s0, &s1);
  fp2_sub(&c2, &c2, &t0); fp2_sub(&c2, &c2, &t2); fp2_add(&c2, &c2, &t1);
  r->c0 = c0; r->c1 = c1; r->c2 = c2;
}
static void fp6_mul_v(fp6 *r, const fp6 *a) {
  fp2 t;
  fp2_mul_xi(&t, &a->c2);
  r->c2 = a->c1;
  r->c1 = a->c0;
  r->c0 = t;
}
static void fp6_inv(fp6 *r, const fp6 *a) {
  fp2 t0, t1, t2, tmp, d;
  fp2_sqr(&t0, &a->c0); fp2_mul(&tmp, &a->c1, &a->c2); fp2_mul_xi(&tmp, &tmp); fp2_sub(&t0, &t0, &tmp);
  fp2_sqr(&t1, &a->c2); fp2_mul_xi(&t1, &t1); fp2_mul(&tmp, &a->c0, &a->c1); fp2_sub(&t1, &t1, &tmp);
  fp2_sqr(&t2, &a->c1); fp2_mul(&tmp, &a->c0, &a->c2); fp2_sub(&t2, &t2, &tmp);
  fp2_mul(&d, &a->c2, &t1); fp2_mul(&tmp, &a->c1, &t2); fp2_add(&d, &d, &tmp); fp2_mul_xi(&d, &d);
  fp2_mul(&tmp, &a->c0, &t0); fp2_add(&d, &d, &tmp);
  fp2_inv(&d, &d);
  fp2_mul(&r->c0, &t0, &d); fp2_mul(&r->c1, &t1, &d); fp2_mul(&r->c2, &t2, &d);
}

static void fp12_one(fp12 *r) { memset(r, 0, sizeof *r); r->c0.c0.c0 = FP_ONE; }
static int fp12_eq(const fp12 *a, const fp12 *b) { return memcmp(a, b, sizeof(fp12)) == 0; }
static void fp12_mul(fp12 *r, const fp12 *a, const fp12 *b) {
  fp6 t0, t1, s0, s1, c1;
  fp6_mul(&t0, &a->c0, &b->c0);
  fp6_mul(&t1, &a->c1, &b->c1);
  fp6_add(&s0, &a->c0, &a->c1);
  fp6_add(&s1, &b->c0, &b->c1);
  fp6_mul(&c1, &s0, &s1);
  fp6_sub(&c1, &c1, &t0);
  fp6_sub(&r->c1, &c1, &t1);
  fp6_mul_v(&t1, &t1);
  fp6_add(&r->c0, &t0, &t1);
}
static void fp12_sqr(fp12 *r, const fp12 *a) { fp12_mul(r, a, a); }
static void fp12_conj(fp12 *r, const fp12 *a) { r->c0 = a->c0; fp6_neg(&r->c1, &a->c1); }
static void fp12_inv(fp12 *r, const fp12 *a) {
  fp6 t0, t1;
  fp6_mul(&t0, &a->c0, &a->c0);
  fp6_mul(&t1, &a->c1, &a->c1);
  fp6_mul_v(&t1, &t1);
  fp6_sub(&t0, &t0, &t1);
  fp6_inv(&t0, &t0);
  fp6_mul(&r->c0, &a->c0, &t0);
  fp6_mul(&t1, &a->c1, &t0);
  fp6_neg(&r->c1, &t1);
}
/* gamma[k] = xi^(k (p-1)/6), k = 0..5: Frobenius on the w-power basis. */
static fp2 FROB_GAMMA[6];
static fp2 *fp12_slot(fp12 *a, int k) {  /* coefficient of w^k */
  fp6 *h = (k & 1) ? &a->c1 : &a->c0;
  int i = k >> 1;
  return i == 0 ? &h->c0 : (i == 1 ? &h->c1 : &h->c2);
}
static void fp12_frob(fp12 *r, const fp12 *a) {
  fp12 t = *a;
  for (int k = 0; k < 6; k++) {
    fp2 *c = fp12_slot(&t, k);
    fp2_conj(c, c);
    fp2_mul(c, c, &FROB_GAMMA[k]);
  }
  *r = t;
}
static void fp12_pow_limbs(fp12 *r, const fp12 *a, const uint64_t *e, int n) {
  fp12 acc, base = *a;
  fp12_one(&acc);
  for (int i = n * 64 - 1; i >= 0; i--) {
    fp12_sqr(&acc, &acc);
    if ((e[i / 64] >> (i % 64)) & 1) fp12_mul(&acc, &acc, &base);
  }
  *r = acc;
}

/* ===================================================================== 4. curves */
typedef struct { fp x, y, z; } g1j;   /* Jacobian, z = 0 is infinity */
typedef struct { fp2 x, y, z; } g2j;

static fp B_G1;   /* 4 */
static fp2 B_G2;  /* 4 (1 + u) */

/* Point arithmetic written once for both groups via a macro over the field. */
#define DEFINE_POINT_OPS(PT, F, F_add, F_sub, F_mul, F_sqr, F_is_zero, F_eq, F_inv, F_neg)           \
  static int PT##_is_inf(const PT *p) { return F_is_zero(&p->z); }                                    \
  static void PT##_dbl(PT *r, const PT *p) {                                                           \
    if (PT##_is_inf(p)) { *r = *p; return; }                                                         \
    F A, B, C, D, E, Fv, t, X3, Y3, Z3;                                                              \
    F_sqr(&A, &p->x); F_sqr(&B, &p->y); F_sqr(&C, &B);                                               \
    F_add(&t, &p->x, &B); F_sqr(&t, &t); F_sub(&t, &t, &A); F_sub(&t, &t, &C); F_add(&D, &t, &t);    \
    F_add(&E, &A, &A); F_add(&E, &E, &A); F_sqr(&Fv, &E);                                            \
    F_sub(&X3, &Fv, &D); F_sub(&X3, &X3, &D);                                                        \
    F_sub(&t, &D, &X3); F_mul(&Y3, &E, &t);                                                          \
    F_add(&C, &C, &C); F_add(&C, &C, &C); F_add(&C, &C, &C); F_sub(&Y3, &Y3, &C);                    \
    F_mul(&Z3, &p->y, &p->z); F_add(&Z3, &Z3, &Z3);                                                  \
    r->x = X3; r->y = Y3; r->z = Z3;                                                                 \
  }                                                                                                  \
  static void PT##_add(PT *r, const PT *p, const PT *q) {                                             \
    if (PT##_is_inf(p)) { *r = *q; return; }                                                         \
    if (PT##_is_inf(q)) { *r = *p; return; }                                                         \
    F Z1Z1, Z2Z2, U1, U2, S1, S2, H, I, J, rr, V, t, X3, Y3, Z3;                                     \
    F_sqr(&Z1Z1, &p->z); F_sqr(&Z2Z2, &q->z);                                                        \
    F_mul(&U1, &p->x, &Z2Z2); F_mul(&U2, &q->x, &Z1Z1);                                              \
    F_mul(&S1, &p->y, &q->z); F_mul(&S1, &S1, &Z2Z2);                                                \
    F_mul(&S2, &q->y, &p->z); F_mul(&S2, &S2, &Z1Z1);                                                \
    F_sub(&H, &U2, &U1); F_sub(&rr, &S2, &S1);                                                       \
    if (F_is_zero(&H)) {                                                                             \
      if (F_is_zero(&rr)) { PT##_dbl(r, p); return; }                                                \
      memset(r, 0, sizeof *r); return;                                                               \
    }                                                                                                \
    F_add(&I, &H, &H); F_sqr(&I, &I); F_mul(&J, &H, &I); F_add(&rr, &rr, &rr);                       \
    F_mul(&V, &U1, &I);                                                                              \
    F_sqr(&X3, &rr); F_sub(&X3, &X3, &J); F_sub(&X3, &X3, &V); F_sub(&X3, &X3, &V);                  \
    F_sub(&t, &V, &X3); F_mul(&Y3, &rr, &t); F_mul(&t, &S1, &J); F_add(&t, &t, &t);                  \
    F_sub(&Y3, &Y3, &t);                                                                             \
    F_add(&t, &p->z, &q->z); F_sqr(&t, &t); F_sub(&t, &t, &Z1Z1); F_sub(&t, &t, &Z2Z2);              \
    F_mul(&Z3, &t, &H);                                                                              \
    r->x = X3; r->y = Y3; r->z = Z3;                                                                 \
  }                                                                                                  \
  static void PT##_neg(PT *r, const PT *p) { *r = *p; F_neg(&r->y, &p->y); }                          \
  /* scalar given as little-endian 64-bit limbs */                                                   \
  static void PT##_mul_limbs(PT *r, const PT *p, const uint64_t *k, int n) {                          \
    PT acc; memset(&acc, 0, sizeof acc);                                                             \
    for (int i = n * 64 - 1; i >= 0; i--) {                                                          \
      PT##_dbl(&acc, &acc);                                                                          \
      if ((k[i / 64] >> (i % 64)) & 1) PT##_add(&acc, &acc, p);                                      \
    }                                                                                                \
    *r = acc;                                                                                        \
  }                                                                                                  \
  static void PT##_to_affine(F *x, F *y, const PT *p) {                                               \
    F zi, zi2, zi3;                                                                                  \
    F_inv(&zi, &p->z); F_sqr(&zi2, &zi); F_mul(&zi3, &zi2, &zi);                                     \
    F_mul(x, &p->x, &zi2); F_mul(y, &p->y, &zi3);                                                    \
  }                                                                                                  \
  static int PT##_eq(const PT *p, const PT *q) {                                                      \
    int pi = PT##_is_inf(p), qi = PT##_is_inf(q);                                                    \
    if (pi || qi) return pi && qi;                                                                   \
    F a, b, c, d; F_sqr(&a, &p->z); F_sqr(&b, &q->z);                                                \
    F_mul(&c, &p->x, &b); F_mul(&d, &q->x, &a);                                                      \
    if (!F_eq(&c, &d)) return 0;                                                                     \
    F_mul(&a, &a, &p->z); F_mul(&b, &b, &q->z);                                                      \
    F_mul(&c, &p->y, &b); F_mul(&d, &q->y, &a);                                                      \
    return F_eq(&c, &d);                                                                             \
  }

DEFINE_POINT_OPS(g1j, fp, fp_add, fp_sub, fp_mul, fp_sqr, fp_is_zero, fp_eq, fp_inv, fp_neg)
DEFINE_POINT_OPS(g2j, fp2, fp2_add, fp2_sub, fp2_mul, fp2_sqr, fp2_is_zero, fp2_eq, fp2_inv, fp2_neg)

static void g1_mul_fr(g1j *r, const g1j *p, const fr *k) {
  uint64_t l[NR];
  fr_to_limbs(l, k);
  g1j_mul_limbs(r, p, l, NR);
}
static void g2_mul_fr(g2j *r, const g2j *p, const fr *k) {
  uint64_t l[NR];
  fr_to_limbs(l, k);
  g2j_mul_limbs(r, p, l, NR);
}
static int g1_on_curve_aff(const fp *x, const fp *y) {
  fp l, rr;
  fp_sqr(&l, y);
  fp_sqr(&rr, x);
  fp_mul(&rr, &rr, x);
  fp_add(&rr, &rr, &B_G1);
  return fp_eq(&l, &rr);
}
static int g1_in_subgroup(const g1j *p) {
  g1j t;
  g1j_mul_limbs(&t, p, R_L, NR);
  return g1j_is_inf(&t);
}
static int g2_in_subgroup(const g2j *p) {
  g2j t;
  g2j_mul_limbs(&t, p, R_L, NR);
  return g2j_is_inf(&t);
}

/* ZCash encoding: bit7 compressed, bit6 infinity, bit5 sign of y. */
static void g1_compress(uint8_t out[48], const g1j *p) {
  if (g1j_is_inf(p)) { memset(out, 0, 48); out[0] = 0xc0; return; }
  fp x, y;
  g1j_to_affine(&x, &y, p);
  fp_to_be(out, &x);
  out[0] |= 0x80;
  if (fp_lex_largest(&y)) out[0] |= 0x20;
}
/* bls12_381 G1Affine::from_compressed: flags, x < p, on curve, torsion free. */
static int g1_decompress(g1j *r, const uint8_t in[48]) {
  uint8_t b[48];
  memcpy(b, in, 48);
  int comp = (b[0] >> 7) & 1, inf = (b[0] >> 6) & 1, sign = (b[0] >> 5) & 1;
  if (!comp) return 0;
  b[0] &= 0x1f;
  if (inf) {
    if (sign) return 0;
    for (int i = 0; i < 48; i++) if (b[i]) return 0;
    memset(r, 0, sizeof *r);
    return 1;
  }
  fp x, y, rhs;
  if (!fp_from_be(&x, b)) return 0;
  fp_sqr(&rhs, &x);
  fp_mul(&rhs, &rhs, &x);
  fp_add(&rhs, &rhs, &B_G1);
  if (!fp_sqrt(&y, &rhs)) return 0;
  if (fp_lex_largest(&y) != sign) fp_neg(&y, &y);
  r->x = x; r->y = y; r->z = FP_ONE;
  return g1_in_subgroup(r);
}
static void g2_compress(uint8_t out[96], const g2j *p) {
  if (g2j_is_inf(p)) { memset(out, 0, 96); out[0] = 0xc0; return; }
  fp2 x, y;
  g2j_to_affine(&x, &y, p);
  fp_to_be(out, &x.c1);
  fp_to_be(out + 48, &x.c0);
  out[0] |= 0x80;
  if (fp2_lex_largest(&y)) out[0] |= 0x20;
}
static int g2_decompress(g2j *r, const uint8_t in[96]) {
  uint8_t b[96];
  memcpy(b, in, 96);
  int comp = (b[0] >> 7) & 1, inf = (b[0] >> 6) & 1, sign = (b[0] >> 5) & 1;
  if (!comp) return 0;
  b[0] &= 0x1f;
  if (inf) {
    if (sign) return 0;
    for (int i = 0; i < 96; i++) if (b[i]) return 0;
    memset(r, 0, sizeof *r);
    return 1;
  }
  fp2 x, y, rhs;
  if (!fp_from_be(&x.c1, b) || !fp_from_be(&x.c0, b + 48)) return 0;
  fp2_sqr(&rhs, &x);
  fp2_mul(&rhs, &rhs, &x);
  fp2_add(&rhs, &rhs, &B_G2);
  if (!fp2_sqrt(&y, &rhs)) return 0;
  if (fp2_lex_largest(&y) != sign) fp2_neg(&y, &y);
  r->x = x; r->y = y; fp2_one(&r->z);
  return g2_in_subgroup(r);
}

static g1j G1_GEN;
static g2j G2_GEN;

/* ===================================================================== 5. pairing */
/* |x| = 0xd201000000010000 ; x is negative. */
static const uint64_t BLS_X_ABS = 0xd201000000010000ULL;

/* Precomputed affine lines of the Miller loop for a G2 point: for every
 * doubling/addition step the slope lambda and c = lambda*x_T - y_T (twist
 * coordinates).  At P = (xp, yp) the line (scaled by w^3) is
 *   c  +  (-lambda*xp) w^2  +  yp w^3 . */
#define MILLER_STEPS 68
typedef struct { fp2 lambda, c; } line_t;
typedef struct { line_t l[MILLER_STEPS]; int n; int inf; } g2_prepared;

static void g2_prepare(g2_prepared *out, const g2j *q) {
  memset(out, 0, sizeof *out);
  if (g2j_is_inf(q)) { out->inf = 1; return; }
  fp2 qx, qy, tx, ty;
  g2j_to_affine(&qx, &qy, q);
  tx = qx; ty = qy;
  int n = 0;
  for (int i = 62; i >= 0; i--) {
    fp2 lam, t, d;
    /* doubling: lambda = 3 tx^2 / (2 ty) */
    fp2_sqr(&t, &tx);
    fp2_add(&lam, &t, &t); fp2_add(&lam, &lam, &t);
    fp2_add(&d, &ty, &ty); fp2_inv(&d, &d);
    fp2_mul(&lam, &lam, &d);
    out->l[n].lambda = lam;
    fp2_mul(&t, &lam, &tx); fp2_sub(&out->l[n].c, &t, &ty);
    n++;
    fp2 nx, ny;
    fp2_sqr(&nx, &lam); fp2_sub(&nx, &nx, &tx); fp2_sub(&nx, &nx, &tx);
    fp2_sub(&t, &tx, &nx); fp2_mul(&ny, &lam, &t); fp2_sub(&ny, &ny, &ty);
    tx = nx; ty = ny;
    if ((BLS_X_ABS >> i) & 1) {
      fp2_sub(&t, &qy, &ty); fp2_sub(&d, &qx, &tx); fp2_inv(&d, &d);
      fp2_mul(&lam, &t, &d);
      out->l[n].lambda = lam;
      fp2_mul(&t, &lam, &tx); fp2_sub(&out->l[n].c, &t, &ty);
      n++;
      fp2_sqr(&nx, &lam); fp2_sub(&nx, &nx, &tx); fp2_sub(&nx, &nx, &qx);
      fp2_sub(&t, &tx, &nx); fp2_mul(&ny, &lam, &t); fp2_sub(&ny, &ny, &ty);
      tx = nx; ty = ny;
    }
  }
  out->n = n;
}

static void fp12_mul_line(fp12 *f, const line_t *ln, const fp *xp, const fp *yp) {
  fp12 l;
  memset(&l, 0, sizeof l);
  l.c0.c0 = ln->c;
  fp2_mul_fp(&l.c0.c1, &ln->lambda, xp);
  fp2_neg(&l.c0.c1, &l.c0.c1);
  l.c1.c1.c0 = *yp;
  fp12_mul(f, f, &l);
}

/* Product of Miller loops f = prod_i f_{|x|,Q_i}(P_i), conjugated for x < 0. */
static void miller_multi(fp12 *f, const g1j *ps, const g2_prepared *qs, int n) {
  fp xs[4], ys[4];
  int skip[4];
  for (int k = 0; k < n; k++) {
    skip[k] = g1j_is_inf(&ps[k]) || qs[k].inf;
    if (!skip[k]) g1j_to_affine(&xs[k], &ys[k], &ps[k]);
  }
  fp12_one(f);
  int idx = 0;
  for (int i = 62; i >= 0; i--) {
    fp12_sqr(f, f);
    for (int k = 0; k < n; k++) if (!skip[k]) fp12_mul_line(f, &qs[k].l[idx], &xs[k], &ys[k]);
    idx++;
    if ((BLS_X_ABS >> i) & 1) {
      for (int k = 0; k < n; k++) if (!skip[k]) fp12_mul_line(f, &qs[k].l[idx], &xs[k], &ys[k]);
      idx++;
    }
  }
  fp12_conj(f, f);
}

static void fp12_exp_by_x(fp12 *r, const fp12 *a) {  /* a^x, x negative; a cyclotomic */
  uint64_t e[1] = {BLS_X_ABS};
  fp12 t;
  fp12_pow_limbs(&t, a, e, 1);
  fp12_conj(r, &t);
}

/* f^((p^12-1)/r * 3) using (p^6-1)(p^2+1) and 3(p^4-p^2+1)/r =
 * (x-1)^2 (x+p) (x^2+p^2-1) + 3. */
static void final_exp(fp12 *r, const fp12 *f) {
  fp12 f1, inv, t, a, b, c, bp2, bc;
  fp12_conj(&f1, f);
  fp12_inv(&inv, f);
  fp12_mul(&f1, &f1, &inv);
  fp12_frob(&t, &f1);
  fp12_frob(&t, &t);
  fp12_mul(&f1, &t, &f1);
  /* a = f1^(x-1) twice */
  fp12_exp_by_x(&a, &f1);
  fp12_conj(&t, &f1);
  fp12_mul(&a, &a, &t);
  fp12_exp_by_x(&t, &a);
  fp12_conj(&b, &a);
  fp12_mul(&a, &t, &b);
  /* b = a^(x+p) */
  fp12_exp_by_x(&b, &a);
  fp12_frob(&t, &a);
  fp12_mul(&b, &b, &t);
  /* c = b^(x^2 + p^2 - 1) */
  fp12_exp_by_x(&c, &b);
  fp12_exp_by_x(&c, &c);
  fp12_frob(&bp2, &b);
  fp12_frob(&bp2, &bp2);
  fp12_mul(&c, &c, &bp2);
  fp12_conj(&bc, &b);
  fp12_mul(&c, &c, &bc);
  /* * f1^3 */
  fp12_sqr(&t, &f1);
  fp12_mul(&t, &t, &f1);
  fp12_mul(r, &c, &t);
}

static uint64_t FE_NAIVE_EXP[72];  /* (p^12 - 1) / r, little-endian limbs */
static int FE_NAIVE_N;

static int fp12_is_one(const fp12 *a) {
  fp12 one;
  fp12_one(&one);
  return fp12_eq(a, &one);
}

/* ===================================================================== 6. hash_to_curve */
static fp ISO_A, ISO_B, SSWU_Z;
static fp ISO_XN[12], ISO_XD[11], ISO_YN[16], ISO_YD[16];

static void fp_horner(fp *r, const fp *c, int n, const fp *x) {
  fp acc = {{0}};
  for (int i = n - 1; i >= 0; i--) {
    fp_mul(&acc, &acc, x);
    fp_add(&acc, &acc, &c[i]);
  }
  *r = acc;
}

/* RFC 9380 6.6.2 simplified SWU onto E', then the 11-isogeny onto E. */
static void map_to_curve_g1(g1j *r, const fp *u) {
  fp tv1, u2, t, x1, gx1, y, x, num;
  fp_sqr(&u2, u);
  fp_mul(&tv1, &u2, &SSWU_Z);        /* Z u^2 */
  fp_sqr(&t, &tv1);                  /* Z^2 u^4 */
  fp_add(&tv1, &t, &tv1);            /* Z^2 u^4 + Z u^2 */
  if (fp_is_zero(&tv1)) {
    fp za;
    fp_mul(&za, &SSWU_Z, &ISO_A);
    fp_inv(&za, &za);
    fp_mul(&x1, &ISO_B, &za);
  } else {
    fp_inv(&tv1, &tv1);
    fp ai, one1;
    fp_add(&one1, &FP_ONE, &tv1);
    fp_inv(&ai, &ISO_A);
    fp_neg(&t, &ISO_B);
    fp_mul(&t, &t, &ai);
    fp_mul(&x1, &t, &one1);
  }
  fp_sqr(&gx1, &x1); fp_add(&gx1, &gx1, &ISO_A); fp_mul(&gx1, &gx1, &x1); fp_add(&gx1, &gx1, &ISO_B);
  if (fp_sqrt(&y, &gx1)) {
    x = x1;
  } else {
    fp gx2;
    fp_mul(&x, &SSWU_Z, &u2);
    fp_mul(&x, &x, &x1);
    fp_sqr(&gx2, &x); fp_add(&gx2, &gx2, &ISO_A); fp_mul(&gx2, &gx2, &x); fp_add(&gx2, &gx2, &ISO_B);
    fp_sqrt(&y, &gx2);
  }
  if (fp_sgn0(u) != fp_sgn0(&y)) fp_neg(&y, &y);
  /* isogeny */
  fp xn, xd, yn, yd;
  fp_horner(&xn, ISO_XN, 12, &x);
  fp_horner(&xd, ISO_XD, 11, &x);
  fp_horner(&yn, ISO_YN, 16, &x);
  fp_horner(&yd, ISO_YD, 16, &x);
  fp_inv(&xd, &xd);
  fp_inv(&yd, &yd);
  fp_mul(&r->x, &xn, &xd);
  fp_mul(&num, &yn, &yd);
  fp_mul(&r->y, &y, &num);
  r->z = FP_ONE;
}

static void hash_to_curve_g1(g1j *r, const uint8_t *msg, size_t len, const uint8_t *dst,
                             size_t dst_len) {
  uint8_t ub[128];
  orc_expand_message_xmd(msg, len, dst, dst_len, ub, 128);
  fp u0, u1;
  fp_from_be_wide(&u0, ub, 64);
  fp_from_be_wide(&u1, ub + 64, 64);
  g1j q0, q1, s;
  map_to_curve_g1(&q0, &u0);
  map_to_curve_g1(&q1, &u1);
  g1j_add(&s, &q0, &q1);
  /* clear_cofactor: h_eff = 1 - x = 0xd201000000010001 */
  uint64_t heff[1] = {0xd201000000010001ULL};
  g1j_mul_limbs(r, &s, heff, 1);
}

/* ===================================================================== init */
static void compute_minv(uint64_t *out, uint64_t m0) {
  uint64_t inv = 1;
  for (int i = 0; i < 7; i++) inv *= 2 - m0 * inv;
  *out = (uint64_t)(0 - inv);
}
static void compute_r2(uint64_t *r2, const uint64_t *m, int n) {
  /* 2^(128 n) mod m by repeated doubling from 1 */
  uint64_t acc[NP] = {1, 0, 0, 0, 0, 0};
  for (int i = 0; i < 128 * n; i++) {
    uint64_t c = mp_add(acc, acc, acc, n);
    if (c || mp_geq(acc, m, n)) mp_sub(acc, acc, m, n);
  }
  memcpy(r2, acc, 8 * n);
}
static void mp_shr(uint64_t *r, const uint64_t *a, int n, int s) {
  for (int i = 0; i < n; i++) {
    uint64_t lo = a[i] >> s;
    uint64_t hi = (i + 1 < n && s) ? a[i + 1] << (64 - s) : 0;
    r[i] = lo | hi;
  }
}

static uint8_t P1_BYTES[48];
static g1j P1_PT;
static const char *P1_HEX =
    "a8ce256102840821a3e94ea9025e4662b205762f9776b3a766c872b948f1fd225e7c59698588e70d11406d161b4e28c9";
static const char *API_ID = "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_";

static void hex_to_bytes(uint8_t *out, const char *hex, size_t n) {
  for (size_t i = 0; i < n; i++) {
    char tmp[3] = {hex[2 * i], hex[2 * i + 1], 0};
    out[i] = (uint8_t)strtoul(tmp, NULL, 16);
  }
}

static pthread_once_t INIT_ONCE = PTHREAD_ONCE_INIT;
static void do_init(void) {
  compute_minv(&P_INV, P_L[0]);
  compute_minv(&R_INV, R_L[0]);
  compute_r2(FP_R2.l, P_L, NP);
  compute_r2(FR_R2.l, R_L, NR);
  {
    fp one = {{1, 0, 0, 0, 0, 0}};
    fp_mul(&FP_ONE, &one, &FP_R2);
    fr onr = {{1, 0, 0, 0}};
    fr_mul(&FR_ONE, &onr, &FR_R2);
  }
  uint64_t two[NP] = {2, 0, 0, 0, 0, 0}, onel[NP] = {1, 0, 0, 0, 0, 0}, three[NP] = {3, 0, 0, 0, 0, 0};
  uint64_t tmp[NP];
  mp_sub(P_MINUS_2, P_L, two, NP);
  mp_add(tmp, P_L, onel, NP);
  mp_shr(P_PLUS1_DIV4, tmp, NP, 2);
  mp_sub(tmp, P_L, onel, NP);
  mp_shr(P_MINUS1_DIV2, tmp, NP, 1);
  mp_sub(tmp, P_L, three, NP);
  mp_shr(P_MINUS3_DIV4, tmp, NP, 2);
  mp_sub(R_MINUS_2, R_L, two, NR);

  fp_from_u64(&B_G1, 4);
  B_G2.c0 = B_G1;
  B_G2.c1 = B_G1;
  /* Frobenius constants: gamma_k = xi^(k (p-1)/6) */
  {
    uint64_t e[NP], pm1[NP];
    mp_sub(pm1, P_L, onel, NP);
    /* divide by 6: shift by 1 then divide by 3 (exact) */
    uint64_t h[NP];
    mp_shr(h, pm1, NP, 1);
    u128 rem = 0;
    for (int i = NP - 1; i >= 0; i--) {
      u128 cur = (rem << 64) | h[i];
      e[i] = (uint64_t)(cur / 3);
      rem = cur % 3;
    }
    fp2 xi = {FP_ONE, FP_ONE}, g;
    fp2_pow(&g, &xi, e, NP);
    fp2_one(&FROB_GAMMA[0]);
    for (int k = 1; k < 6; k++) fp2_mul(&FROB_GAMMA[k], &FROB_GAMMA[k - 1], &g);
  }
  /* (p^12 - 1) / r for the naive final exponentiation check */
  {
    uint64_t pw[72] = {0}, acc[72] = {0};
    memcpy(pw, P_L, sizeof P_L);
    memcpy(acc, P_L, sizeof P_L);
    for (int k = 1; k < 12; k++) {
      uint64_t prod[72] = {0};
      for (int i = 0; i < 72; i++) {
        if (!acc[i]) continue;
        u128 c = 0;
        for (int j = 0; j < NP && i + j < 72; j++) {
          c += (u128)acc[i] * P_L[j] + prod[i + j];
          prod[i + j] = (uint64_t)c;
          c >>= 64;
        }
        for (int j = i + NP; j < 72 && c; j++) {
          c += prod[j];
          prod[j] = (uint64_t)c;
          c >>= 64;
        }
      }
      memcpy(acc, prod, sizeof acc);
    }
    uint64_t one72[72] = {1};
    mp_sub(acc, acc, one72, 72);
    /* long division by r (4 limbs) via bitwise shift-subtract */
    uint64_t q[72] = {0}, rem[5] = {0};
    for (int i = 72 * 64 - 1; i >= 0; i--) {
      /* rem = rem*2 + bit */
      uint64_t carry = (acc[i / 64] >> (i % 64)) & 1;
      for (int j = 0; j < 5; j++) {
        uint64_t nc = rem[j] >> 63;
        rem[j] = (rem[j] << 1) | carry;
        carry = nc;
      }
      uint64_t r5[5] = {R_L[0], R_L[1], R_L[2], R_L[3], 0};
      if (mp_geq(rem, r5, 5)) {
        mp_sub(rem, rem, r5, 5);
        q[i / 64] |= 1ULL << (i % 64);
      }
    }
    memcpy(FE_NAIVE_EXP, q, sizeof q);
    FE_NAIVE_N = 72;
    while (FE_NAIVE_N > 1 && FE_NAIVE_EXP[FE_NAIVE_N - 1] == 0) FE_NAIVE_N--;
  }
  /* hash_to_curve constants */
  fp_from_hex(&ISO_A, H2C_ISO_A);
  fp_from_hex(&ISO_B, H2C_ISO_B);
  fp_from_u64(&SSWU_Z, H2C_SSWU_Z);
  for (int i = 0; i < 12; i++) fp_from_hex(&ISO_XN[i], H2C_ISO_XNUM[i]);
  for (int i = 0; i < 11; i++) fp_from_hex(&ISO_XD[i], H2C_ISO_XDEN[i]);
  for (int i = 0; i < 16; i++) fp_from_hex(&ISO_YN[i], H2C_ISO_YNUM[i]);
  for (int i = 0; i < 16; i++) fp_from_hex(&ISO_YD[i], H2C_ISO_YDEN[i]);
  /* generators of G1 and G2 (standard BLS12-381 generators, compressed form) */
  {
    uint8_t b[96];
    hex_to_bytes(b, "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb", 48);
    g1_decompress(&G1_GEN, b);
    hex_to_bytes(b, "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8", 96);
    g2_decompress(&G2_GEN, b);
  }
  hex_to_bytes(P1_BYTES, P1_HEX, 48);
  g1_decompress(&P1_PT, P1_BYTES);
}
int orc_init(void) {
  pthread_once(&INIT_ONCE, do_init);
  return 0;
}

/* ===================================================================== exported primitives */
int orc_hash_to_scalar(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                       uint8_t out[32]) {
  /* crates/cubbs-bbs-protocol/src/utils/util.rs:104-133 */
  orc_init();
  uint8_t ub[48];
  if (orc_expand_message_xmd(msg, len, dst, dst_len, ub, 48)) return ORC_ERR_INPUT;
  fr s;
  fr_from_be_wide(&s, ub, 48);
  fr_to_be(out, &s);
  return ORC_OK;
}

static void api_dst(uint8_t *out, size_t *len, const char *suffix) {
  size_t a = strlen(API_ID), s = strlen(suffix);
  memcpy(out, API_ID, a);
  memcpy(out + a, suffix, s);
  *len = a + s;
}

int orc_map_messages_to_scalars(const uint8_t *msgs, const size_t *lens, size_t count,
                                uint8_t *out) {
  /* crates/cubbs-bbs-protocol/src/utils/message.rs:47-64 */
  uint8_t dst[128];
  size_t dl;
  api_dst(dst, &dl, "MAP_MSG_TO_SCALAR_AS_HASH_");
  size_t off = 0;
  for (size_t i = 0; i < count; i++) {
    orc_hash_to_scalar(msgs + off, lens[i], dst, dl, out + 32 * i);
    off += lens[i];
  }
  return ORC_OK;
}

int orc_seeded_random_scalars(const uint8_t *seed, size_t seed_len, const uint8_t *dst,
                              size_t dst_len, size_t count, uint8_t *out) {
  /* crates/cubbs-bbs-protocol/src/utils/util.rs:330-357 */
  orc_init();
  size_t n = 48 * count;
  uint8_t *v = (uint8_t *)malloc(n ? n : 1);
  if (orc_expand_message_xmd(seed, seed_len, dst, dst_len, v, n)) { free(v); return ORC_ERR_INPUT; }
  for (size_t i = 0; i < count; i++) {
    fr s;
    fr_from_be_wide(&s, v + 48 * i, 48);
    fr_to_be(out + 32 * i, &s);
  }
  free(v);
  return ORC_OK;
}

/* Seeded blinding convention of this project: ChaCha20 (RFC 8439) keyed by the
 * 32-byte seed, zero nonce, counter from 0; scalar k is the 64-byte keystream
 * chunk k read as a little-endian integer, reduced mod r. */
int orc_chacha_random_scalars(const uint8_t seed[32], size_t count, uint8_t *out) {
  orc_init();
  uint8_t nonce[12] = {0};
  for (size_t i = 0; i < count; i++) {
    uint8_t blk[64], be[64];
    orc_chacha20_block(seed, (uint32_t)i, nonce, blk);
    for (int k = 0; k < 64; k++) be[k] = blk[63 - k];
    fr s;
    fr_from_be_wide(&s, be, 64);
    fr_to_be(out + 32 * i, &s);
  }
  return ORC_OK;
}

int orc_g1_decompress_check(const uint8_t in[48]) {
  orc_init();
  g1j p;
  return g1_decompress(&p, in);
}
int orc_g2_decompress_check(const uint8_t in[96]) {
  orc_init();
  g2j p;
  return g2_decompress(&p, in);
}
int orc_g1_add(const uint8_t a[48], const uint8_t b[48], uint8_t out[48]) {
  orc_init();
  g1j p, q, s;
  if (!g1_decompress(&p, a) || !g1_decompress(&q, b)) return ORC_ERR_INPUT;
  g1j_add(&s, &p, &q);
  g1_compress(out, &s);
  return ORC_OK;
}
int orc_g1_mul(const uint8_t p[48], const uint8_t k[32], uint8_t out[48]) {
  orc_init();
  g1j a, s;
  fr kk;
  if (!g1_decompress(&a, p) || !fr_from_be(&kk, k)) return ORC_ERR_INPUT;
  g1_mul_fr(&s, &a, &kk);
  g1_compress(out, &s);
  return ORC_OK;
}
int orc_g1_msm(const uint8_t *points, const uint8_t *scalars, size_t n, uint8_t out[48]) {
  orc_init();
  g1j acc;
  memset(&acc, 0, sizeof acc);
  for (size_t i = 0; i < n; i++) {
    g1j a, s;
    fr kk;
    if (!g1_decompress(&a, points + 48 * i) || !fr_from_be(&kk, scalars + 32 * i)) return ORC_ERR_INPUT;
    g1_mul_fr(&s, &a, &kk);
    g1j_add(&acc, &acc, &s);
  }
  g1_compress(out, &acc);
  return ORC_OK;
}
int orc_g1_generator(uint8_t out[48]) { orc_init(); g1_compress(out, &G1_GEN); return ORC_OK; }
int orc_g2_generator(uint8_t out[96]) { orc_init(); g2_compress(out, &G2_GEN); return ORC_OK; }
int orc_g2_mul(const uint8_t q[96], const uint8_t k[32], uint8_t out[96]) {
  orc_init();
  g2j a, s;
  fr kk;
  if (!g2_decompress(&a, q) || !fr_from_be(&kk, k)) return ORC_ERR_INPUT;
  g2_mul_fr(&s, &a, &kk);
  g2_compress(out, &s);
  return ORC_OK;
}
int orc_hash_to_curve_g1(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                         uint8_t out[48]) {
  orc_init();
  g1j p;
  hash_to_curve_g1(&p, msg, len, dst, dst_len);
  g1_compress(out, &p);
  return ORC_OK;
}

static int pairing_product_is_one(const g1j *ps, const g2j *qs, int n) {
  g2_prepared prep[4];
  for (int i = 0; i < n; i++) g2_prepare(&prep[i], &qs[i]);
  fp12 f, r;
  miller_multi(&f, ps, prep, n);
  final_exp(&r, &f);
  return fp12_is_one(&r);
}

int orc_pairing_eq(const uint8_t a1[48], const uint8_t a2[96], const uint8_t b1[48],
                   const uint8_t b2[96]) {
  orc_init();
  g1j ps[2];
  g2j qs[2];
  if (!g1_decompress(&ps[0], a1) || !g2_decompress(&qs[0], a2) || !g1_decompress(&ps[1], b1) ||
      !g2_decompress(&qs[1], b2))
    return ORC_ERR_INPUT;
  g1j_neg(&ps[1], &ps[1]);
  return pairing_product_is_one(ps, qs, 2);
}

int orc_pairing_gt(const uint8_t p[48], const uint8_t q[96], int naive, uint8_t out[576]) {
  orc_init();
  g1j a;
  g2j b;
  if (!g1_decompress(&a, p) || !g2_decompress(&b, q)) return ORC_ERR_INPUT;
  g2_prepared prep;
  g2_prepare(&prep, &b);
  fp12 f, r;
  miller_multi(&f, &a, &prep, 1);
  if (naive)
    fp12_pow_limbs(&r, &f, FE_NAIVE_EXP, FE_NAIVE_N);
  else
    final_exp(&r, &f);
  for (int k = 0; k < 6; k++) {
    fp12 *rr = &r;
    const fp6 *h = (k < 3) ? &rr->c0 : &rr->c1;
    const fp2 *c = (k % 3 == 0) ? &h->c0 : (k % 3 == 1 ? &h->c1 : &h->c2);
    fp_to_be(out + 96 * k, &c->c0);
    fp_to_be(out + 96 * k + 48, &c->c1);
  }
  return ORC_OK;
}

/* Self-check of the final exponentiation: the x-chain result must equal the
 * cube of the plain (p^12 - 1)/r power (3 (p^4 - p^2 + 1)/r =
 * (x-1)^2 (x+p) (x^2+p^2-1) + 3).  Returns 1 when it holds. */
int orc_check_final_exp(const uint8_t p[48], const uint8_t q[96]) {
  orc_init();
  g1j a;
  g2j b;
  if (!g1_decompress(&a, p) || !g2_decompress(&b, q)) return ORC_ERR_INPUT;
  g2_prepared prep;
  g2_prepare(&prep, &b);
  fp12 f, fast, naive, cube;
  miller_multi(&f, &a, &prep, 1);
  final_exp(&fast, &f);
  fp12_pow_limbs(&naive, &f, FE_NAIVE_EXP, FE_NAIVE_N);
  fp12_sqr(&cube, &naive);
  fp12_mul(&cube, &cube, &naive);
  return fp12_eq(&fast, &cube) && !fp12_is_one(&fast);
}

int orc_fr_mul(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
  orc_init();
  fr x, y;
  if (!fr_from_be(&x, a) || !fr_from_be(&y, b)) return ORC_ERR_INPUT;
  fr_mul(&x, &x, &y);
  fr_to_be(out, &x);
  return ORC_OK;
}
int orc_fr_add(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]) {
  orc_init();
  fr x, y;
  if (!fr_from_be(&x, a) || !fr_from_be(&y, b)) return ORC_ERR_INPUT;
  fr_add(&x, &x, &y);
  fr_to_be(out, &x);
  return ORC_OK;
}
int orc_fr_inv(const uint8_t a[32], uint8_t out[32]) {
  orc_init();
  fr x;
  if (!fr_from_be(&x, a) || fr_is_zero(&x)) return ORC_ERR_INPUT;
  fr_inv(&x, &x);
  fr_to_be(out, &x);
  return ORC_OK;
}

/* ===================================================================== 7. BBS */
/* Generator cache: Q1, H_1, ... (crates/cubbs-bbs-protocol/src/bbsplus/generators.rs:119-148).
 * It only grows, under GEN_LOCK.  A growth copies the entries into new
 * arrays and then publishes them; the previous arrays are kept, never freed,
 * so a thread that is reading entries it made sure of (ensure_generators)
 * through the pointers it loaded before the growth keeps reading valid
 * memory.  Readers load the pointers with gen_pts() / gen_comp(). */
static pthread_mutex_t GEN_LOCK = PTHREAD_MUTEX_INITIALIZER;
static g1j *GEN_PTS = NULL;
static uint8_t *GEN_COMP = NULL;
static size_t GEN_COUNT = 0, GEN_CAP = 0;
static uint8_t GEN_V[48];

static const g1j *gen_pts(void) { return __atomic_load_n(&GEN_PTS, __ATOMIC_ACQUIRE); }
static const uint8_t *gen_comp(void) { return __atomic_load_n(&GEN_COMP, __ATOMIC_ACQUIRE); }

static int ensure_generators(size_t count) {
  pthread_mutex_lock(&GEN_LOCK);
  if (GEN_COUNT >= count) { pthread_mutex_unlock(&GEN_LOCK); return 0; }
  if (count > GEN_CAP) {
    size_t cap = GEN_CAP ? GEN_CAP : 64;
    while (cap < count) cap *= 2;
    g1j *pts = (g1j *)malloc(cap * sizeof(g1j));
    uint8_t *comp = (uint8_t *)malloc(cap * 48);
    if (GEN_COUNT) {
      memcpy(pts, GEN_PTS, GEN_COUNT * sizeof(g1j));
      memcpy(comp, GEN_COMP, GEN_COUNT * 48);
    }
    __atomic_store_n(&GEN_PTS, pts, __ATOMIC_RELEASE);
    __atomic_store_n(&GEN_COMP, comp, __ATOMIC_RELEASE);
    GEN_CAP = cap;
  }
  uint8_t seed_dst[128], gen_dst[128], seed[128];
  size_t sdl, gdl, sl;
  api_dst(seed_dst, &sdl, "SIG_GENERATOR_SEED_");
  api_dst(gen_dst, &gdl, "SIG_GENERATOR_DST_");
  if (GEN_COUNT == 0) {
    api_dst(seed, &sl, "MESSAGE_GENERATOR_SEED");
    orc_expand_message_xmd(seed, sl, seed_dst, sdl, GEN_V, 48);
  }
  for (size_t i = GEN_COUNT + 1; i <= count; i++) {
    uint8_t buf[56];
    memcpy(buf, GEN_V, 48);
    for (int k = 0; k < 8; k++) buf[48 + k] = (uint8_t)((uint64_t)i >> (56 - 8 * k));
    orc_expand_message_xmd(buf, 56, seed_dst, sdl, GEN_V, 48);
    hash_to_curve_g1(&GEN_PTS[i - 1], GEN_V, 48, gen_dst, gdl);
    g1_compress(GEN_COMP + 48 * (i - 1), &GEN_PTS[i - 1]);
  }
  GEN_COUNT = count;
  pthread_mutex_unlock(&GEN_LOCK);
  return 0;
}

int orc_generators(size_t count, uint8_t *out) {
  orc_init();
  ensure_generators(count);
  memcpy(out, gen_comp(), 48 * count);
  return ORC_OK;
}
int orc_p1(uint8_t out[48]) { orc_init(); memcpy(out, P1_BYTES, 48); return ORC_OK; }

/* calculate_domain: crates/cubbs-bbs-protocol/src/utils/util.rs:147-195 */
static void calc_domain(fr *out, const uint8_t pk[96], size_t L, const uint8_t *header,
                        size_t hlen) {
  ensure_generators(L + 1);
  size_t api_len = strlen(API_ID);
  size_t n = 96 + 8 + 48 * (L + 1) + api_len + 8 + hlen;
  uint8_t *buf = (uint8_t *)malloc(n);
  size_t o = 0;
  memcpy(buf + o, pk, 96); o += 96;
  for (int k = 0; k < 8; k++) buf[o + k] = (uint8_t)((uint64_t)L >> (56 - 8 * k));
  o += 8;
  memcpy(buf + o, gen_comp(), 48 * (L + 1)); o += 48 * (L + 1);
  memcpy(buf + o, API_ID, api_len); o += api_len;
  for (int k = 0; k < 8; k++) buf[o + k] = (uint8_t)((uint64_t)hlen >> (56 - 8 * k));
  o += 8;
  if (hlen) memcpy(buf + o, header, hlen);
  o += hlen;
  uint8_t dst[128], s[32];
  size_t dl;
  api_dst(dst, &dl, "H2S_");
  orc_hash_to_scalar(buf, o, dst, dl, s);
  fr_from_be(out, s);
  free(buf);
}

int orc_domain(const uint8_t pk[96], size_t L, const uint8_t *header, size_t header_len,
               uint8_t out[32]) {
  orc_init();
  fr d;
  calc_domain(&d, pk, L, header, header_len);
  fr_to_be(out, &d);
  return ORC_OK;
}

static void msg_scalars(fr *out, const uint8_t *msgs, const size_t *lens, size_t L) {
  uint8_t dst[128], s[32];
  size_t dl;
  api_dst(dst, &dl, "MAP_MSG_TO_SCALAR_AS_HASH_");
  size_t off = 0;
  for (size_t i = 0; i < L; i++) {
    orc_hash_to_scalar(msgs + off, lens[i], dst, dl, s);
    fr_from_be(&out[i], s);
    off += lens[i];
  }
}

/* B = P1 + Q1 * domain + sum H_i * m_i  (signature.rs:620-640) */
static void compute_b(g1j *B, const fr *domain, const fr *m, size_t L) {
  g1j t;
  g1_mul_fr(&t, &gen_pts()[0], domain);
  g1j_add(B, &P1_PT, &t);
  for (size_t i = 0; i < L; i++) {
    g1_mul_fr(&t, &gen_pts()[i + 1], &m[i]);
    g1j_add(B, B, &t);
  }
}

int orc_compute_b(const uint8_t pk[96], const uint8_t *header, size_t header_len,
                  const uint8_t *msgs, const size_t *lens, size_t L, uint8_t out[48]) {
  orc_init();
  fr dom, *m = (fr *)malloc(sizeof(fr) * (L ? L : 1));
  calc_domain(&dom, pk, L, header, header_len);
  msg_scalars(m, msgs, lens, L);
  g1j B;
  compute_b(&B, &dom, m, L);
  g1_compress(out, &B);
  free(m);
  return ORC_OK;
}

int orc_keygen(const uint8_t *ikm, size_t ikm_len, const uint8_t *info, size_t info_len,
               uint8_t sk[32], uint8_t pk[96]) {
  /* crates/cubbs-bbs-protocol/src/bbsplus/keys.rs:214-260 */
  orc_init();
  if (ikm_len < 32 || info_len > 65535) return ORC_ERR_INPUT;
  uint8_t *buf = (uint8_t *)malloc(ikm_len + 2 + info_len);
  memcpy(buf, ikm, ikm_len);
  buf[ikm_len] = (uint8_t)(info_len >> 8);
  buf[ikm_len + 1] = (uint8_t)info_len;
  if (info_len) memcpy(buf + ikm_len + 2, info, info_len);
  uint8_t dst[128];
  size_t dl;
  api_dst(dst, &dl, "KEYGEN_DST_");
  orc_hash_to_scalar(buf, ikm_len + 2 + info_len, dst, dl, sk);
  free(buf);
  return orc_sk_to_pk(sk, pk);
}
int orc_sk_to_pk(const uint8_t sk[32], uint8_t pk[96]) {
  orc_init();
  fr s;
  if (!fr_from_be(&s, sk)) return ORC_ERR_INPUT;
  g2j w;
  g2_mul_fr(&w, &G2_GEN, &s);
  g2_compress(pk, &w);
  return ORC_OK;
}

/* core_sign: crates/cubbs-bbs-protocol/src/bbsplus/signature.rs:544-580,642-674 */
int orc_sign(const uint8_t sk[32], const uint8_t pk[96], const uint8_t *header, size_t header_len,
             const uint8_t *msgs, const size_t *lens, size_t L, uint8_t sig[80]) {
  orc_init();
  fr skf;
  g2j pkp;
  if (!fr_from_be(&skf, sk)) return ORC_ERR_INPUT;
  if (!g2_decompress(&pkp, pk)) return ORC_ERR_INPUT;
  fr dom, *m = (fr *)malloc(sizeof(fr) * (L ? L : 1));
  calc_domain(&dom, pk, L, header, header_len);
  msg_scalars(m, msgs, lens, L);
  /* e = hash_to_scalar(sk || m_1..m_L || domain, api_id || "H2S_") */
  uint8_t *buf = (uint8_t *)malloc(32 * (L + 2));
  fr_to_be(buf, &skf);
  for (size_t i = 0; i < L; i++) fr_to_be(buf + 32 * (i + 1), &m[i]);
  fr_to_be(buf + 32 * (L + 1), &dom);
  uint8_t dst[128], eb[32];
  size_t dl;
  api_dst(dst, &dl, "H2S_");
  orc_hash_to_scalar(buf, 32 * (L + 2), dst, dl, eb);
  free(buf);
  fr e;
  fr_from_be(&e, eb);
  g1j B, A;
  compute_b(&B, &dom, m, L);
  free(m);
  fr se;
  fr_add(&se, &skf, &e);
  if (fr_is_zero(&se)) return ORC_ERR_CRYPTO;
  fr_inv(&se, &se);
  g1_mul_fr(&A, &B, &se);
  if (g1j_is_inf(&A)) return ORC_ERR_CRYPTO;
  g1_compress(sig, &A);
  memcpy(sig + 48, eb, 32);
  return ORC_OK;
}

/* crates/cubbs-bbs/src/protocol.rs:318-326 (length and parse errors) and
 * signature.rs:676-728: e(A, PK + e*BP2) * e(B, -BP2) == 1. */
int orc_verify(const uint8_t *pk, size_t pk_len, const uint8_t *sig, size_t sig_len,
               const uint8_t *header, size_t header_len, const uint8_t *msgs, const size_t *lens,
               size_t L) {
  orc_init();
  if (pk_len != 96 || sig_len != 80) return ORC_ERR_INPUT;
  g2j W;
  g1j A;
  fr e;
  if (!g2_decompress(&W, pk)) return ORC_ERR_INPUT;
  if (!g1_decompress(&A, sig) || !fr_from_be(&e, sig + 48)) return ORC_ERR_INPUT;
  fr dom, *m = (fr *)malloc(sizeof(fr) * (L ? L : 1));
  calc_domain(&dom, pk, L, header, header_len);
  msg_scalars(m, msgs, lens, L);
  g1j B;
  compute_b(&B, &dom, m, L);
  free(m);
  g2j A2, t;
  g2_mul_fr(&t, &G2_GEN, &e);
  g2j_add(&A2, &W, &t);
  g1j ps[2] = {A, B};
  g2j qs[2];
  qs[0] = A2;
  g2j_neg(&qs[1], &G2_GEN);
  return pairing_product_is_one(ps, qs, 2);
}

/* core_proof_gen / proof_init / proof_challenge / proof_finalize:
 * crates/cubbs-bbs-protocol/src/bbsplus/proof.rs:1568-1645,1754-1815,2155-2297,2317-2377 */
static void challenge_hash(fr *c, const uint8_t comp[5][48], const fr *dom, const size_t *idx,
                           const fr *dm, size_t R, const uint8_t *ph, size_t ph_len) {
  size_t n = 8 + 40 * R + 240 + 32 + 8 + ph_len;
  uint8_t *buf = (uint8_t *)malloc(n);
  size_t o = 0;
  for (int k = 0; k < 8; k++) buf[o + k] = (uint8_t)((uint64_t)R >> (56 - 8 * k));
  o += 8;
  for (size_t i = 0; i < R; i++) {
    for (int k = 0; k < 8; k++) buf[o + k] = (uint8_t)((uint64_t)idx[i] >> (56 - 8 * k));
    o += 8;
    fr_to_be(buf + o, &dm[i]);
    o += 32;
  }
  for (int k = 0; k < 5; k++) { memcpy(buf + o, comp[k], 48); o += 48; }
  fr_to_be(buf + o, dom); o += 32;
  for (int k = 0; k < 8; k++) buf[o + k] = (uint8_t)((uint64_t)ph_len >> (56 - 8 * k));
  o += 8;
  if (ph_len) memcpy(buf + o, ph, ph_len);
  o += ph_len;
  uint8_t dst[128], s[32];
  size_t dl;
  api_dst(dst, &dl, "H2S_");
  orc_hash_to_scalar(buf, o, dst, dl, s);
  fr_from_be(c, s);
  free(buf);
}

long orc_proof_gen(const uint8_t pk[96], const uint8_t *sig, size_t sig_len, const uint8_t *header,
                   size_t header_len, const uint8_t *ph, size_t ph_len, const uint8_t *msgs,
                   const size_t *lens, size_t L, const size_t *disclosed, size_t R,
                   const uint8_t *random_scalars, uint8_t *proof_out, size_t proof_cap,
                   uint8_t *trace_out) {
  orc_init();
  /* crates/cubbs-bbs/src/protocol.rs:423-437: strictly increasing indexes < L */
  for (size_t i = 0; i < R; i++) {
    if (disclosed[i] >= L || (i && disclosed[i] <= disclosed[i - 1])) return ORC_ERR_INPUT;
  }
  g2j W;
  if (!g2_decompress(&W, pk)) return ORC_ERR_INPUT;
  if (sig_len != 80) return ORC_ERR_INPUT;
  g1j A;
  fr e;
  if (!g1_decompress(&A, sig) || !fr_from_be(&e, sig + 48)) return ORC_ERR_INPUT;
  size_t U = L - R;
  size_t need = 272 + 32 * U;
  if (proof_cap < need) return ORC_ERR_INPUT;
  fr *rs = (fr *)malloc(sizeof(fr) * (5 + U));
  for (size_t i = 0; i < 5 + U; i++)
    if (!fr_from_be(&rs[i], random_scalars + 32 * i)) { free(rs); return ORC_ERR_INPUT; }
  fr dom, *m = (fr *)malloc(sizeof(fr) * (L ? L : 1));
  calc_domain(&dom, pk, L, header, header_len);
  msg_scalars(m, msgs, lens, L);
  g1j B;
  compute_b(&B, &dom, m, L);
  const fr *r1 = &rs[0], *r2 = &rs[1], *et = &rs[2], *r1t = &rs[3], *r3t = &rs[4];
  g1j D, Abar, Bbar, T1, T2, t, t2;
  fr r1r2;
  fr_mul(&r1r2, r1, r2);
  g1_mul_fr(&D, &B, r2);
  g1_mul_fr(&Abar, &A, &r1r2);
  g1_mul_fr(&t, &D, r1);
  g1_mul_fr(&t2, &Abar, &e);
  g1j_neg(&t2, &t2);
  g1j_add(&Bbar, &t, &t2);
  g1_mul_fr(&t, &Abar, et);
  g1_mul_fr(&t2, &D, r1t);
  g1j_add(&T1, &t, &t2);
  g1_mul_fr(&T2, &D, r3t);
  size_t *und = (size_t *)malloc(sizeof(size_t) * (U ? U : 1));
  size_t u = 0, d = 0;
  for (size_t i = 0; i < L; i++) {
    if (d < R && disclosed[d] == i) { d++; continue; }
    und[u++] = i;
  }
  for (size_t j = 0; j < U; j++) {
    g1_mul_fr(&t, &gen_pts()[und[j] + 1], &rs[5 + j]);
    g1j_add(&T2, &T2, &t);
  }
  uint8_t comp[5][48];
  g1_compress(comp[0], &Abar);
  g1_compress(comp[1], &Bbar);
  g1_compress(comp[2], &D);
  g1_compress(comp[3], &T1);
  g1_compress(comp[4], &T2);
  fr *dm = (fr *)malloc(sizeof(fr) * (R ? R : 1));
  for (size_t i = 0; i < R; i++) dm[i] = m[disclosed[i]];
  fr c;
  challenge_hash(&c, comp, &dom, disclosed, dm, R, ph, ph_len);
  /* finalize */
  fr r3, ecap, r1cap, r3cap, tmp;
  if (fr_is_zero(r2)) { free(m); free(und); free(dm); free(rs); return ORC_ERR_CRYPTO; }
  fr_inv(&r3, r2);
  fr_mul(&tmp, &e, &c); fr_add(&ecap, et, &tmp);
  fr_mul(&tmp, r1, &c); fr_sub(&r1cap, r1t, &tmp);
  fr_mul(&tmp, &r3, &c); fr_sub(&r3cap, r3t, &tmp);
  memcpy(proof_out, comp[0], 48);
  memcpy(proof_out + 48, comp[1], 48);
  memcpy(proof_out + 96, comp[2], 48);
  fr_to_be(proof_out + 144, &ecap);
  fr_to_be(proof_out + 176, &r1cap);
  fr_to_be(proof_out + 208, &r3cap);
  for (size_t j = 0; j < U; j++) {
    fr mc;
    fr_mul(&tmp, &m[und[j]], &c);
    fr_add(&mc, &rs[5 + j], &tmp);
    fr_to_be(proof_out + 240 + 32 * j, &mc);
  }
  fr_to_be(proof_out + 240 + 32 * U, &c);
  if (trace_out) {
    for (int k = 0; k < 5; k++) memcpy(trace_out + 48 * k, comp[k], 48);
    fr_to_be(trace_out + 240, &dom);
    fr_to_be(trace_out + 272, &c);
  }
  free(m); free(und); free(dm); free(rs);
  return (long)need;
}

/* crates/cubbs-bbs/src/protocol.rs:772-793 and proof.rs:189-225,2379-2426,3015-3080 */
int orc_proof_verify(const uint8_t *pk, size_t pk_len, const uint8_t *proof, size_t proof_len,
                     const uint8_t *header, size_t header_len, const uint8_t *ph, size_t ph_len,
                     const uint8_t *disclosed_msgs, const size_t *lens, size_t msg_count,
                     const size_t *disclosed, size_t R, size_t total_message_count) {
  orc_init();
  if (msg_count != R) return 0;
  for (size_t i = 0; i < R; i++) {
    if (disclosed[i] >= total_message_count || (i && disclosed[i] <= disclosed[i - 1])) return 0;
  }
  if (pk_len != 96) return ORC_ERR_INPUT;
  g2j W;
  if (!g2_decompress(&W, pk)) return ORC_ERR_INPUT;
  /* BBSplusPoKSignature::from_bytes: 3 points, 3 scalars, then 32-byte chunks
   * (a trailing partial chunk is ignored), the last chunk is the challenge. */
  if (proof_len < 240 + 32) return ORC_ERR_INPUT;
  g1j Abar, Bbar, D;
  if (!g1_decompress(&Abar, proof) || !g1_decompress(&Bbar, proof + 48) ||
      !g1_decompress(&D, proof + 96))
    return ORC_ERR_INPUT;
  fr ecap, r1cap, r3cap, c;
  if (!fr_from_be(&ecap, proof + 144) || !fr_from_be(&r1cap, proof + 176) ||
      !fr_from_be(&r3cap, proof + 208))
    return ORC_ERR_INPUT;
  size_t chunks = (proof_len - 240) / 32;
  size_t U = chunks - 1;
  fr *mcap = (fr *)malloc(sizeof(fr) * (chunks ? chunks : 1));
  for (size_t j = 0; j < chunks; j++)
    if (!fr_from_be(&mcap[j], proof + 240 + 32 * j)) { free(mcap); return ORC_ERR_INPUT; }
  c = mcap[U];
  size_t L = U + R;
  for (size_t i = 0; i < R; i++) if (disclosed[i] > L - 1) { free(mcap); return 0; }
  fr dom;
  calc_domain(&dom, pk, L, header, header_len);
  fr *dm = (fr *)malloc(sizeof(fr) * (R ? R : 1));
  msg_scalars(dm, disclosed_msgs, lens, R);
  /* T1 = Bbar*c + Abar*e^ + D*r1^ */
  g1j T1, T2, t, Bv;
  g1_mul_fr(&T1, &Bbar, &c);
  g1_mul_fr(&t, &Abar, &ecap); g1j_add(&T1, &T1, &t);
  g1_mul_fr(&t, &D, &r1cap); g1j_add(&T1, &T1, &t);
  /* Bv = P1 + Q1*dom + sum_R H_i m_i ; T2 = Bv*c + D*r3^ + sum_U H_j m^_j */
  g1_mul_fr(&t, &gen_pts()[0], &dom);
  g1j_add(&Bv, &P1_PT, &t);
  for (size_t i = 0; i < R; i++) {
    g1_mul_fr(&t, &gen_pts()[disclosed[i] + 1], &dm[i]);
    g1j_add(&Bv, &Bv, &t);
  }
  g1_mul_fr(&T2, &Bv, &c);
  g1_mul_fr(&t, &D, &r3cap); g1j_add(&T2, &T2, &t);
  size_t u = 0, d = 0;
  for (size_t i = 0; i < L; i++) {
    if (d < R && disclosed[d] == i) { d++; continue; }
    g1_mul_fr(&t, &gen_pts()[i + 1], &mcap[u]);
    g1j_add(&T2, &T2, &t);
    u++;
  }
  uint8_t comp[5][48];
  memcpy(comp[0], proof, 48);
  memcpy(comp[1], proof + 48, 48);
  memcpy(comp[2], proof + 96, 48);
  /* re-serialise parsed points exactly as the reference does (to_affine().to_compressed()) */
  g1_compress(comp[0], &Abar);
  g1_compress(comp[1], &Bbar);
  g1_compress(comp[2], &D);
  g1_compress(comp[3], &T1);
  g1_compress(comp[4], &T2);
  fr cc;
  challenge_hash(&cc, comp, &dom, disclosed, dm, R, ph, ph_len);
  free(mcap);
  free(dm);
  if (!fr_eq(&cc, &c)) return 0;
  g1j ps[2] = {Abar, Bbar};
  g2j qs[2];
  qs[0] = W;
  g2j_neg(&qs[1], &G2_GEN);
  return pairing_product_is_one(ps, qs, 2);
}
