// SHA-256 (FIPS 180-4) and RFC 9380 expand_message_xmd on the device.
//
// One lane hashes one input.  Two forms:
//  * block-aligned inputs laid out by the planner (h2s_blocks, the message
//    hashes): one loop over all blocks with the compression inlined once, so
//    the state and the block stay in VGPRs and nothing goes to scratch;
//  * a byte stream (Sha256: domains, transcripts, generator seeds): bytes are
//    appended through a 32-bit shift register and committed to the block with
//    a select (no dynamically indexed register arrays).  Its helpers and the
//    compression are out of line and take the state by reference, so that
//    state lives in the lane's scratch memory; these inputs are few.
// Inputs are described as a sequence of byte sources (global memory, constant
// strings, values already in registers).
#pragma once
#include "field.hpp"

__constant__ static const uint32_t SHA256_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

DEV uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// The 64 rounds, inlined where the state and the block are locals indexed by
// constants only, so both stay in VGPRs.
DEV void sha256_rounds(uint32_t h[8], uint32_t w[16]) {
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    uint32_t wi;
    if (i < 16) {
      wi = w[i];
    } else {
      uint32_t w15 = w[(i - 15) & 15], w2 = w[(i - 2) & 15];
      uint32_t s0 = rotr32(w15, 7) ^ rotr32(w15, 18) ^ (w15 >> 3);
      uint32_t s1 = rotr32(w2, 17) ^ rotr32(w2, 19) ^ (w2 >> 10);
      wi = w[i & 15] + s0 + w[(i - 7) & 15] + s1;
      w[i & 15] = wi;
    }
    uint32_t S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = hh + S1 + ch + SHA256_K[i] + wi;
    uint32_t S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    hh = g;
    g = f;
    f = e;
    e = d + t1;
    d = c;
    c = b;
    b = a;
    a = t1 + mj + S0;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d;
  h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}
// One out-of-line copy for the byte-stream hasher below, whose state lives
// in memory across its helpers.
DEVN void sha256_compress(uint32_t h[8], uint32_t w[16]) { sha256_rounds(h, w); }

struct Sha256 {
  uint32_t h[8];
  uint32_t w[16];
  uint32_t cur;    // partial word being assembled
  uint32_t n;      // bytes absorbed (including any prefix blocks)
};

DEV void sha_init(Sha256& s) {
  s.h[0] = 0x6a09e667; s.h[1] = 0xbb67ae85; s.h[2] = 0x3c6ef372; s.h[3] = 0xa54ff53a;
  s.h[4] = 0x510e527f; s.h[5] = 0x9b05688c; s.h[6] = 0x1f83d9ab; s.h[7] = 0x5be0cd19;
  s.cur = 0;
  s.n = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) s.w[i] = 0;
}

// State after one all-zero 64-byte block (the Z_pad prefix of expand_message_xmd).
DEV void sha_init_after_zpad(Sha256& s) {
  sha_init(s);
  sha256_compress(s.h, s.w);
#pragma unroll
  for (int i = 0; i < 16; i++) s.w[i] = 0;
  s.n = 64;
}

DEVI void sha_byte(Sha256& s, uint32_t b) {
  s.cur = (s.cur << 8) | (b & 0xff);
  const uint32_t pos = s.n & 63;
  s.n++;
  if ((pos & 3) == 3) {
    const uint32_t wi = pos >> 2;
#pragma unroll
    for (int k = 0; k < 16; k++)
      if ((uint32_t)k == wi) s.w[k] = s.cur;
    if (wi == 15) sha256_compress(s.h, s.w);
  }
}

// Fast path for a 32-bit big-endian word when the stream is word aligned.
DEVI void sha_word(Sha256& s, uint32_t wv) {
  if ((s.n & 3) == 0) {
    const uint32_t wi = (s.n & 63) >> 2;
    s.n += 4;
#pragma unroll
    for (int k = 0; k < 16; k++)
      if ((uint32_t)k == wi) s.w[k] = wv;
    if (wi == 15) sha256_compress(s.h, s.w);
  } else {
    sha_byte(s, wv >> 24);
    sha_byte(s, wv >> 16);
    sha_byte(s, wv >> 8);
    sha_byte(s, wv);
  }
}

// Big-endian word of the 4 bytes at q[0] >> sh (sh = 8 * (address & 3)),
// taking the high bytes from the next aligned dword when unaligned.
DEV uint32_t sha_be_word(uint32_t lo, uint32_t hi, uint32_t sh) {
  return __builtin_bswap32((uint32_t)((((uint64_t)hi << 32) | lo) >> sh));
}

// Bytes from global memory.  After aligning the stream to a word, bytes are
// read as aligned dwords (funnel-shifted when the source is unaligned; the
// next dword is read only when it holds bytes of the range), and whole 64-byte
// blocks go straight into the block registers from 16-17 loads issued together.
DEVI void sha_bytes_global(Sha256& s, const uint8_t* p, uint32_t len) {
  uint32_t i = 0;
  while (i < len && (s.n & 3)) sha_byte(s, p[i++]);
  const uintptr_t a = (uintptr_t)(p + i);
  const uint32_t* q = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
  const uint32_t sh = (uint32_t)(a & 3) * 8;
  while (i + 4 <= len && (s.n & 63)) {
    const uint32_t lo = q[0], hi = sh ? q[1] : 0u;
    sha_word(s, sha_be_word(lo, hi, sh));
    q++;
    i += 4;
  }
  while (i + 64 <= len) {
    uint32_t d[17];
#pragma unroll
    for (int k = 0; k < 16; k++) d[k] = q[k];
    d[16] = sh ? q[16] : 0u;
#pragma unroll
    for (int k = 0; k < 16; k++) s.w[k] = sha_be_word(d[k], d[k + 1], sh);
    s.n += 64;
    sha256_compress(s.h, s.w);
    q += 16;
    i += 64;
  }
  while (i + 4 <= len) {
    const uint32_t lo = q[0], hi = sh ? q[1] : 0u;
    sha_word(s, sha_be_word(lo, hi, sh));
    q++;
    i += 4;
  }
  while (i < len) sha_byte(s, p[i++]);
}

// Bytes of a constant string (a DST, the API id): the same word-level path.
DEVI void sha_bytes_const(Sha256& s, const uint8_t* p, uint32_t len) { sha_bytes_global(s, p, len); }

DEV void sha_u64_be(Sha256& s, uint64_t v) {
  sha_word(s, (uint32_t)(v >> 32));
  sha_word(s, (uint32_t)v);
}

// Finish: padding and length; digest words left in s.h (big-endian words).
// After the 0x80 byte and the zero bytes that complete its word, the rest of
// the block is zeroed as words (the compression leaves its schedule in w, so
// unwritten words are not zero); the 64-bit length takes words 14 and 15,
// in one more block when fewer than two words are left.
DEVN void sha_final(Sha256& s) {
  const uint64_t bits = (uint64_t)s.n * 8;
  sha_byte(s, 0x80);
  while (s.n & 3) sha_byte(s, 0);
  const uint32_t wi = (s.n & 63) >> 2;  // next free word of the block
#pragma unroll
  for (int k = 0; k < 16; k++)
    if ((uint32_t)k >= wi) s.w[k] = 0;
  if (wi > 14) {
    sha256_compress(s.h, s.w);
#pragma unroll
    for (int k = 0; k < 14; k++) s.w[k] = 0;
  }
  s.w[14] = (uint32_t)(bits >> 32);
  s.w[15] = (uint32_t)bits;
  sha256_compress(s.h, s.w);
}

// ------------------------------------------------------------------ domain-separation tags
// API_ID = "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_" (44 bytes)
// DST strings are API_ID || suffix, followed in DST_prime by their length.
struct Dst {
  const uint8_t* bytes;
  uint32_t len;
};

__constant__ static const uint8_t DST_MAP_MSG[] =
    "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_MAP_MSG_TO_SCALAR_AS_HASH_";
__constant__ static const uint8_t DST_H2S[] = "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_H2S_";
__constant__ static const uint8_t API_ID_BYTES[] = "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_";
__constant__ static const uint8_t DST_GEN_SEED[] =
    "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_SIG_GENERATOR_SEED_";
__constant__ static const uint8_t DST_GEN[] =
    "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_SIG_GENERATOR_DST_";
#define API_ID_LEN 44u
#define DST_MAP_MSG_LEN 70u
#define DST_H2S_LEN 48u
#define DST_GEN_SEED_LEN 63u
#define DST_GEN_LEN 62u

// Append l_i_b_str || 0x00 || DST || len(DST) after the message of b_0.
DEVN void xmd_b0_suffix(Sha256& s, uint32_t out_len, const uint8_t* dst, uint32_t dst_len) {
  sha_byte(s, out_len >> 8);
  sha_byte(s, out_len);
  sha_byte(s, 0);
  sha_bytes_const(s, dst, dst_len);
  sha_byte(s, dst_len);
  sha_final(s);
}

// b_i = H(prev_xor || i || DST_prime) where prev_xor is 8 words.
DEVN void xmd_bi(uint32_t out[8], const uint32_t in[8], uint32_t idx, const uint8_t* dst,
                 uint32_t dst_len) {
  Sha256 s;
  sha_init(s);
#pragma unroll
  for (int k = 0; k < 8; k++) sha_word(s, in[k]);
  sha_byte(s, idx);
  sha_bytes_const(s, dst, dst_len);
  sha_byte(s, dst_len);
  sha_final(s);
#pragma unroll
  for (int k = 0; k < 8; k++) out[k] = s.h[k];
}

// Given the finished b_0 state, produce 48 bytes (b_1 || b_2[0..16]) as
// 12 big-endian words.
DEVN void xmd48_from_b0(uint32_t okm[12], const uint32_t b0[8], const uint8_t* dst, uint32_t dst_len) {
  uint32_t b1[8], x[8], b2[8];
  xmd_bi(b1, b0, 1, dst, dst_len);
#pragma unroll
  for (int k = 0; k < 8; k++) x[k] = b0[k] ^ b1[k];
  xmd_bi(b2, x, 2, dst, dst_len);
#pragma unroll
  for (int k = 0; k < 8; k++) okm[k] = b1[k];
#pragma unroll
  for (int k = 0; k < 4; k++) okm[8 + k] = b2[k];
}

// OS2IP(48 bytes as 12 big-endian words) mod r -> canonical scalar limbs.
DEVN void scalar_from_okm48(Fr& out, const uint32_t okm[12]) {
  uint32_t v[16];
#pragma unroll
  for (int k = 0; k < 12; k++) v[k] = okm[11 - k];
#pragma unroll
  for (int k = 12; k < 16; k++) v[k] = 0;
  fr_reduce_wide(out, v);
}

// hash_to_scalar of a message laid out by the planner (crates/cubbs-bbs-
// protocol/src/utils/util.rs:104-133): the input of b_0 after its all-zero
// Z_pad block, i.e. msg || I2OSP(48, 2) || 0x00 || DST' and the SHA-256
// padding, as `blocks` blocks of big-endian words (16-byte aligned); `zpad`
// is the state after the Z_pad block; `tail` the constant words of the b_1
// and b_2 inputs (tail_blocks blocks each, b_1's first, words 0..7 of each
// one's first block left for b_0 and b_0 ^ b_1).  All blocks run through one
// loop with one inlined compression, so the state, the block and b_0, b_1
// stay in registers; every lane of a wave with the same block count runs
// the same compressions.
DEV Fr h2s_blocks(const uint32_t* words, uint32_t blocks, const uint32_t* zpad,
                  const uint32_t* tail, uint32_t tail_blocks) {
  uint32_t h[8], b0[8], b1[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    h[k] = zpad[k];
    b0[k] = 0;
    b1[k] = 0;
  }
  const uint32_t n1 = blocks + tail_blocks, n = n1 + tail_blocks;
#pragma unroll 1
  for (uint32_t b = 0; b < n; b++) {
    const uint32_t* src = b < blocks ? words + 16 * (size_t)b : tail + 16 * (size_t)(b - blocks);
    const uint4* q = reinterpret_cast<const uint4*>(src);
    uint32_t w[16];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint4 v = q[k];
      w[4 * k] = v.x;
      w[4 * k + 1] = v.y;
      w[4 * k + 2] = v.z;
      w[4 * k + 3] = v.w;
    }
    if (b == blocks || b == n1) {
      // first block of b_1 = H(b_0 || 1 || DST') or b_2 = H((b_0 ^ b_1) || 2 || DST')
      const bool second = b == n1;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (second) {
          b1[k] = h[k];
          w[k] = b0[k] ^ h[k];
        } else {
          b0[k] = h[k];
          w[k] = h[k];
        }
      }
      h[0] = 0x6a09e667; h[1] = 0xbb67ae85; h[2] = 0x3c6ef372; h[3] = 0xa54ff53a;
      h[4] = 0x510e527f; h[5] = 0x9b05688c; h[6] = 0x1f83d9ab; h[7] = 0x5be0cd19;
    }
    sha256_rounds(h, w);
  }
  // okm = b_1 || b_2[0..16] as 12 big-endian words; OS2IP mod r
  uint32_t v[16];
#pragma unroll
  for (int k = 0; k < 8; k++) v[11 - k] = b1[k];
#pragma unroll
  for (int k = 0; k < 4; k++) v[3 - k] = h[k];
#pragma unroll
  for (int k = 12; k < 16; k++) v[k] = 0;
  Fr out;
  fr_reduce_wide(out, v);
  return out;
}

// hash_to_scalar(msg, dst) for a message in global memory (crates/cubbs-bbs-
// protocol/src/utils/util.rs:104-133): expand_message_xmd(msg, dst, 48),
// OS2IP mod r.
DEVN void hash_to_scalar_global(Fr& out, const uint8_t* msg, uint32_t len, const uint8_t* dst,
                                uint32_t dst_len) {
  Sha256 s;
  sha_init_after_zpad(s);
  sha_bytes_global(s, msg, len);
  xmd_b0_suffix(s, 48, dst, dst_len);
  uint32_t okm[12];
  xmd48_from_b0(okm, s.h, dst, dst_len);
  scalar_from_okm48(out, okm);
}
