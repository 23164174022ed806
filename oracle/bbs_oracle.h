/*
 * CPU oracle for the BBS (BLS12-381, SHA-256) hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark call it.  The HIP
 * product never links, loads or calls it.
 *
 * It restates, in plain C with 64-bit limbs, the algorithms of the reference
 * protocol crate:
 *   crates/cubbs-bbs-protocol/src/bbsplus/{ciphersuites,generators,keys,
 *     signature,proof}.rs, src/utils/{util,message}.rs
 * and the third-party arithmetic that crate calls (bls12_381_plus 0.8.18:
 * field/curve/pairing/codec; elliptic-curve 0.13.8 ExpandMsgXmd; sha2 0.10.9),
 * restated from RFC 9380, FIPS 180-4 and the BLS12-381 definition.
 *
 * Conventions: scalars are 32-byte big-endian, G1 points 48-byte compressed,
 * G2 points 96-byte compressed (ZCash encoding).  Messages are passed packed:
 * one byte buffer plus an array of lengths.
 * Return codes: 1 = valid/true, 0 = invalid/false, negative = error.
 */
#ifndef BBS_ORACLE_H
#define BBS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_OK 0
#define ORC_ERR_INPUT (-1)
#define ORC_ERR_CRYPTO (-2)

int orc_init(void);

/* primitives */
void orc_sha256(const uint8_t *msg, size_t len, uint8_t out[32]);
int orc_expand_message_xmd(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                           uint8_t *out, size_t out_len);
int orc_hash_to_scalar(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                       uint8_t out[32]);
int orc_map_messages_to_scalars(const uint8_t *msgs, const size_t *lens, size_t count,
                                uint8_t *out /* 32*count */);
int orc_seeded_random_scalars(const uint8_t *seed, size_t seed_len, const uint8_t *dst,
                              size_t dst_len, size_t count, uint8_t *out);
int orc_chacha_random_scalars(const uint8_t seed[32], size_t count, uint8_t *out);
void orc_chacha20_block(const uint8_t key[32], uint32_t counter, const uint8_t nonce[12],
                        uint8_t out[64]);

/* curve / codec */
int orc_g1_decompress_check(const uint8_t in[48]);            /* 1 valid, 0 invalid */
int orc_g2_decompress_check(const uint8_t in[96]);
int orc_g1_add(const uint8_t a[48], const uint8_t b[48], uint8_t out[48]);
int orc_g1_mul(const uint8_t p[48], const uint8_t k[32], uint8_t out[48]);
int orc_g1_msm(const uint8_t *points, const uint8_t *scalars, size_t n, uint8_t out[48]);
int orc_g1_generator(uint8_t out[48]);
int orc_g2_generator(uint8_t out[96]);
int orc_g2_mul(const uint8_t q[96], const uint8_t k[32], uint8_t out[96]);
int orc_hash_to_curve_g1(const uint8_t *msg, size_t len, const uint8_t *dst, size_t dst_len,
                         uint8_t out[48]);
/* e(a1, a2) == e(b1, b2) ? 1 : 0 */
int orc_pairing_eq(const uint8_t a1[48], const uint8_t a2[96], const uint8_t b1[48],
                   const uint8_t b2[96]);
/* GT value of e(P,Q) after final exponentiation, 12 Fp coefficients (c0.c0.c0 ...),
 * each 48-byte big-endian.  `naive` selects the plain (p^12-1)/r exponent. */
int orc_pairing_gt(const uint8_t p[48], const uint8_t q[96], int naive, uint8_t out[576]);

int orc_check_final_exp(const uint8_t p[48], const uint8_t q[96]);

/* scalar field helpers (canonical check, arithmetic) */
int orc_fr_mul(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
int orc_fr_add(const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
int orc_fr_inv(const uint8_t a[32], uint8_t out[32]);

/* BBS */
int orc_keygen(const uint8_t *ikm, size_t ikm_len, const uint8_t *info, size_t info_len,
               uint8_t sk[32], uint8_t pk[96]);
int orc_sk_to_pk(const uint8_t sk[32], uint8_t pk[96]);
int orc_generators(size_t count, uint8_t *out /* 48*count: Q1, H_1.. */);
int orc_p1(uint8_t out[48]);
int orc_domain(const uint8_t pk[96], size_t L, const uint8_t *header, size_t header_len,
               uint8_t out[32]);
int orc_compute_b(const uint8_t pk[96], const uint8_t *header, size_t header_len,
                  const uint8_t *msgs, const size_t *lens, size_t L, uint8_t out[48]);
int orc_sign(const uint8_t sk[32], const uint8_t pk[96], const uint8_t *header, size_t header_len,
             const uint8_t *msgs, const size_t *lens, size_t L, uint8_t sig[80]);
int orc_verify(const uint8_t *pk, size_t pk_len, const uint8_t *sig, size_t sig_len,
               const uint8_t *header, size_t header_len, const uint8_t *msgs, const size_t *lens,
               size_t L);
/* random_scalars: (5 + U) * 32 bytes [r1, r2, e~, r1~, r3~, m~_1..m~_U].
 * trace_out (optional, may be NULL): Abar|Bbar|D|T1|T2 (5*48) | domain | challenge. */
long orc_proof_gen(const uint8_t pk[96], const uint8_t *sig, size_t sig_len, const uint8_t *header,
                   size_t header_len, const uint8_t *ph, size_t ph_len, const uint8_t *msgs,
                   const size_t *lens, size_t L, const size_t *disclosed, size_t R,
                   const uint8_t *random_scalars, uint8_t *proof_out, size_t proof_cap,
                   uint8_t *trace_out);
int orc_proof_verify(const uint8_t *pk, size_t pk_len, const uint8_t *proof, size_t proof_len,
                     const uint8_t *header, size_t header_len, const uint8_t *ph, size_t ph_len,
                     const uint8_t *disclosed_msgs, const size_t *lens, size_t msg_count,
                     const size_t *disclosed, size_t R, size_t total_message_count);

#ifdef __cplusplus
}
#endif
#endif
