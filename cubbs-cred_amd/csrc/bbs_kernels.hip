// Device kernels of the BBS hot path (gfx950).  One lane owns one job
// (message, term chunk, signature, pairing check); every kernel is bounds-
// checked against its job count and exits on all lanes.
#include "bbs_internal.h"
#include "dev_io.hpp"
#include "h2c.hpp"
#include "pairing12.hpp"

namespace cubbs {

#define GRID(n, b) dim3(((n) + (b) - 1) / (b)), dim3(b)

// ------------------------------------------------------------------ hash to scalar
__global__ __launch_bounds__(256) void k_h2s(const H2sJob* __restrict__ jobs, uint32_t n,
                                            const uint8_t* __restrict__ blob,
                                            const uint8_t* __restrict__ dst, uint32_t dst_len,
                                            uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  H2sJob j = jobs[i];
  Fr s;
  hash_to_scalar_global(s, blob + j.off, j.len, dst, dst_len);
  store_fr(out + (size_t)j.out_index * FR_WORDS, s);
}
void launch_h2s(hipStream_t s, const H2sJob* jobs, uint32_t n, const uint8_t* blob,
                const uint8_t* dst, uint32_t dst_len, uint32_t* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_h2s, GRID(n, 64), 0, s, jobs, n, blob, dst, dst_len, out);
}

// Packed hash jobs with a caller DST in device memory: kind 0 = SHA-256
// digest (32 bytes), 1 = expand_message_xmd(msg, dst, 48), 2 = hash_to_scalar.
__global__ __launch_bounds__(256) void k_hash_jobs(int kind, const H2sJob* __restrict__ jobs,
                                                  uint32_t n, const uint8_t* __restrict__ blob,
                                                  const uint8_t* __restrict__ dst, uint32_t dst_len,
                                                  uint8_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  H2sJob j = jobs[i];
  if (kind == 0) {
    Sha256 s;
    sha_init(s);
    sha_bytes_global(s, blob + j.off, j.len);
    sha_final(s);
    for (int k = 0; k < 8; k++) {
      uint32_t w = s.h[k];
      uint8_t* o = out + (size_t)32 * j.out_index + 4 * k;
      o[0] = w >> 24; o[1] = w >> 16; o[2] = w >> 8; o[3] = w;
    }
    return;
  }
  Sha256 s;
  sha_init_after_zpad(s);
  sha_bytes_global(s, blob + j.off, j.len);
  xmd_b0_suffix(s, 48, dst, dst_len);
  uint32_t okm[12];
  xmd48_from_b0(okm, s.h, dst, dst_len);
  if (kind == 1) {
    for (int k = 0; k < 12; k++) {
      uint32_t w = okm[k];
      uint8_t* o = out + (size_t)48 * j.out_index + 4 * k;
      o[0] = w >> 24; o[1] = w >> 16; o[2] = w >> 8; o[3] = w;
    }
    return;
  }
  Fr sc;
  scalar_from_okm48(sc, okm);
  fr_to_be_bytes(out + (size_t)32 * j.out_index, sc);
}
void launch_hash_jobs(hipStream_t s, int kind, const H2sJob* jobs, uint32_t n, const uint8_t* blob,
                      const uint8_t* dst, uint32_t dst_len, uint8_t* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_hash_jobs, GRID(n, 64), 0, s, kind, jobs, n, blob, dst, dst_len, out);
}

// domain = hash_to_scalar(PK || I2OSP(L, 8) || Q1 || H_1..H_L || api_id ||
//                         I2OSP(len(header), 8) || header, api_id || "H2S_")
DEVN void domain_job(const DomainJob& j, const uint8_t* hdr_blob, const uint8_t* pk_bytes,
                     const uint8_t* gen_comp, uint32_t* out) {
  Sha256 s;
  sha_init_after_zpad(s);
  sha_bytes_global(s, pk_bytes + (size_t)j.pk_slot * 96, 96);
  sha_u64_be(s, j.L);
  sha_bytes_global(s, gen_comp + 48 * (size_t)GEN_Q1, 48u * (j.L + 1));
  sha_bytes_const(s, API_ID_BYTES, API_ID_LEN);
  sha_u64_be(s, j.hdr_len);
  sha_bytes_global(s, hdr_blob + j.hdr_off, j.hdr_len);
  xmd_b0_suffix(s, 48, DST_H2S, DST_H2S_LEN);
  uint32_t okm[12];
  xmd48_from_b0(okm, s.h, DST_H2S, DST_H2S_LEN);
  Fr d;
  scalar_from_okm48(d, okm);
  store_fr(out + (size_t)j.out_index * FR_WORDS, d);
}
__global__ __launch_bounds__(256) void k_domain(const DomainJob* __restrict__ jobs, uint32_t n,
                                               const uint8_t* __restrict__ hdr_blob,
                                               const uint8_t* __restrict__ pk_bytes,
                                               const uint8_t* __restrict__ gen_comp,
                                               uint32_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  domain_job(jobs[i], hdr_blob, pk_bytes, gen_comp, out);
}
// Message hashes and domains are independent: one launch runs both, the long
// domain jobs on lanes [0, n_dom), the message jobs on [n_dom, n_dom + n_msg)
// (ordered by block count, so the lanes of a wave compress in step).
__global__ __launch_bounds__(256) void k_hash_stage(const DomainJob* __restrict__ dom, uint32_t n_dom,
                                                   const MsgHashJob* __restrict__ msg, uint32_t n_msg,
                                                   const uint8_t* __restrict__ blob,
                                                   const uint32_t* __restrict__ zpad,
                                                   const uint32_t* __restrict__ tail, uint32_t tail_blocks,
                                                   const uint8_t* __restrict__ pk_bytes,
                                                   const uint8_t* __restrict__ gen_comp,
                                                   uint32_t* __restrict__ out) {
  wave_priority<PRIO_EARLY>();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_dom) {
    domain_job(dom[i], blob, pk_bytes, gen_comp, out);
  } else if (i < n_dom + n_msg) {
    const MsgHashJob j = msg[i - n_dom];
    const Fr sc = h2s_blocks(reinterpret_cast<const uint32_t*>(blob + j.off), j.blocks, zpad, tail,
                             tail_blocks);
    store_fr(out + (size_t)j.out_index * FR_WORDS, sc);
  }
}
void launch_hash_stage(hipStream_t s, const DomainJob* dom, uint32_t n_dom, const MsgHashJob* msg,
                       uint32_t n_msg, const uint8_t* blob, const uint32_t* zpad, const uint32_t* tail,
                       uint32_t tail_blocks, const uint8_t* pk_bytes, const uint8_t* gen_comp,
                       uint32_t* out) {
  const uint32_t n = n_dom + n_msg;
  if (!n) return;
  hipLaunchKernelGGL(k_hash_stage, GRID(n, 64), 0, s, dom, n_dom, msg, n_msg, blob, zpad, tail,
                     tail_blocks, pk_bytes, gen_comp, out);
}
__global__ void k_sha_zpad(uint32_t* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  Sha256 st;
  sha_init_after_zpad(st);
  for (int k = 0; k < 8; k++) out[k] = st.h[k];
}
void launch_sha_zpad(hipStream_t s, uint32_t* out) {
  hipLaunchKernelGGL(k_sha_zpad, dim3(1), dim3(1), 0, s, out);
}
void launch_domain(hipStream_t s, const DomainJob* jobs, uint32_t n, const uint8_t* hdr_blob,
                   const uint8_t* pk_bytes, const uint8_t* gen_comp, uint32_t* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_domain, GRID(n, 64), 0, s, jobs, n, hdr_blob, pk_bytes, gen_comp, out);
}

// ------------------------------------------------------------------ generators
// Seed chain (generators.rs:119-148): v = expand(api_id || "MESSAGE_GENERATOR_SEED",
// seed_dst, 48); v_i = expand(v_{i-1} || I2OSP(i, 8), seed_dst, 48).  Sequential:
// one lane.  v_state holds the last v (12 words) and is updated.
__global__ void k_gen_seeds(uint32_t first, uint32_t count, uint32_t* v_state, uint32_t* seeds) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint32_t v[12];
  if (first == 1) {
    Sha256 s;
    sha_init_after_zpad(s);
    sha_bytes_const(s, API_ID_BYTES, API_ID_LEN);
    const char* tag = "MESSAGE_GENERATOR_SEED";
    for (int k = 0; k < 22; k++) sha_byte(s, (uint8_t)tag[k]);
    xmd_b0_suffix(s, 48, DST_GEN_SEED, DST_GEN_SEED_LEN);
    xmd48_from_b0(v, s.h, DST_GEN_SEED, DST_GEN_SEED_LEN);
  } else {
    for (int k = 0; k < 12; k++) v[k] = v_state[k];
  }
  for (uint32_t i = first; i < first + count; i++) {
    Sha256 s;
    sha_init_after_zpad(s);
    for (int k = 0; k < 12; k++) sha_word(s, v[k]);
    sha_u64_be(s, i);
    xmd_b0_suffix(s, 48, DST_GEN_SEED, DST_GEN_SEED_LEN);
    xmd48_from_b0(v, s.h, DST_GEN_SEED, DST_GEN_SEED_LEN);
    for (int k = 0; k < 12; k++) seeds[(size_t)(i - first) * 12 + k] = v[k];
  }
  for (int k = 0; k < 12; k++) v_state[k] = v[k];
}
void launch_gen_seeds(hipStream_t s, uint32_t first, uint32_t count, uint32_t* v_state,
                      uint32_t* seeds) {
  hipLaunchKernelGGL(k_gen_seeds, dim3(1), dim3(1), 0, s, first, count, v_state, seeds);
}

// Generator from chain seed v_i (i >= 1) = hash_to_curve(v_i, api_id ||
// "SIG_GENERATOR_DST_"); it occupies slot i (slot 1 = Q1, slot 0 = P1).
__global__ __launch_bounds__(64) void k_gen_points(uint32_t first, uint32_t count,
                                                  const uint32_t* __restrict__ seeds,
                                                  uint32_t* __restrict__ gen_affine,
                                                  uint8_t* __restrict__ gen_comp) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t v[12];
  for (int k = 0; k < 12; k++) v[k] = seeds[(size_t)t * 12 + k];
  G1J p;
  hash_to_curve_g1_words(p, v, 12, DST_GEN, DST_GEN_LEN);
  G1A a;
  pt_to_affine(a, p);
  uint32_t slot = first + t;
  store_g1a(gen_affine + (size_t)slot * G1A_WORDS, a);
  uint8_t c[48];
  g1_compress_affine(c, a);
  for (int k = 0; k < 48; k++) gen_comp[(size_t)slot * 48 + k] = c[k];
  if (first == 1 && t == 0) {
    // slot 0: P1 (ciphersuite constant)
    G1A p1;
    g1_load_affine_const(p1, bls::BBS_P1);
    store_g1a(gen_affine, p1);
    g1_compress_affine(c, p1);
    for (int k = 0; k < 48; k++) gen_comp[k] = c[k];
  }
}
void launch_gen_points(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* seeds,
                       uint32_t* gen_affine, uint8_t* gen_comp) {
  if (!count) return;
  hipLaunchKernelGGL(k_gen_points, GRID(count, 64), 0, s, first, count, seeds, gen_affine, gen_comp);
}

// Comb tables: one lane per (generator, window): base = 2^(8w) G, entries
// d * base for d = 1..128 built with mixed additions in Jacobian form, then
// normalised to affine with one inversion (Montgomery's batch trick).
// `scratch` holds 24 words per entry and lane (z_d and the prefix product).
__global__ __launch_bounds__(64) void k_comb_build(uint32_t first, uint32_t count,
                                                  const uint32_t* __restrict__ gen_affine,
                                                  uint32_t* __restrict__ tables,
                                                  uint32_t* __restrict__ scratch) {
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * COMB_WINDOWS) return;
  uint32_t g = first + t / COMB_WINDOWS;
  uint32_t w = t % COMB_WINDOWS;
  G1A ga;
  load_g1a(ga, gen_affine + (size_t)g * G1A_WORDS);
  G1J base;
  pt_from_affine(base, ga);
  for (uint32_t k = 0; k < 8 * w; k++) pt_dbl(base, base);
  G1A baseA;
  pt_to_affine(baseA, base);
  uint32_t* out = tables + ((size_t)g * COMB_WINDOWS + w) * COMB_ENTRIES * COMB_ENTRY_WORDS;
  uint32_t* sc = scratch + (size_t)t * COMB_ENTRIES * COMB_ENTRY_WORDS;
  G1J acc;
  pt_from_affine(acc, baseA);
  Fp prod;
  fp_one(prod);
  for (int d = 1; d <= COMB_ENTRIES; d++) {
    fp_mul(prod, prod, acc.z);
    uint32_t* e = out + (size_t)(d - 1) * COMB_ENTRY_WORDS;
    uint32_t* z = sc + (size_t)(d - 1) * COMB_ENTRY_WORDS;
    for (int k = 0; k < FPW; k++) {
      e[k] = acc.x.v[k];
      e[FPW + k] = acc.y.v[k];
      z[k] = acc.z.v[k];
      z[FPW + k] = prod.v[k];
    }
    if (d < COMB_ENTRIES) pt_madd(acc, acc, baseA);
  }
  Fp inv;
  fp_inv(inv, prod);  // 1 / (z_1 ... z_128)
  for (int d = COMB_ENTRIES; d >= 1; d--) {
    uint32_t* e = out + (size_t)(d - 1) * COMB_ENTRY_WORDS;
    const uint32_t* z = sc + (size_t)(d - 1) * COMB_ENTRY_WORDS;
    Fp zd, pre, zi;
    for (int k = 0; k < FPW; k++) zd.v[k] = z[k];
    if (d > 1) {
      for (int k = 0; k < FPW; k++) pre.v[k] = sc[(size_t)(d - 2) * COMB_ENTRY_WORDS + FPW + k];
      fp_mul(zi, inv, pre);
    } else {
      zi = inv;
    }
    fp_mul(inv, inv, zd);
    Fp X, Y, zi2, zi3;
    for (int k = 0; k < FPW; k++) {
      X.v[k] = e[k];
      Y.v[k] = e[FPW + k];
    }
    fp_sqr(zi2, zi);
    fp_mul(zi3, zi2, zi);
    G1A a;
    fp_mul(a.x, X, zi2);
    fp_mul(a.y, Y, zi3);
    store_g1a(e, a);
  }
}
void launch_comb_build(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* gen_affine,
                       uint32_t* tables, uint32_t* scratch) {
  if (!count) return;
  uint32_t n = count * COMB_WINDOWS;
  hipLaunchKernelGGL(k_comb_build, GRID(n, 64), 0, s, first, count, gen_affine, tables, scratch);
}

// Wide comb tables: one lane per (generator, window, run of COMB16_RUN
// entries).  base = 2^(16w) G is entry d = 1 of window 2w of the generator's
// 8-bit table; the run starts at d0 * base (short double-and-add), continues
// by mixed additions in Jacobian form and is normalised to affine with one
// inversion (Montgomery's batch trick).  `scratch` holds z_d and the running
// product, 28 words per entry and lane.
__global__ __launch_bounds__(64) void k_comb16_build(uint32_t first, uint32_t count,
                                                    const uint32_t* __restrict__ tables,
                                                    uint32_t* __restrict__ tables16,
                                                    uint32_t* __restrict__ scratch) {
  constexpr uint32_t RUNS = COMB16_ENTRIES / COMB16_RUN;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * COMB16_WINDOWS * RUNS) return;
  const uint32_t g = first + t / (COMB16_WINDOWS * RUNS);
  const uint32_t w = (t / RUNS) % COMB16_WINDOWS;
  const uint32_t d0 = (t % RUNS) * COMB16_RUN + 1;
  G1A base;
  load_g1a(base, tables + ((size_t)g * COMB_WINDOWS + 2 * w) * COMB_ENTRIES * COMB_ENTRY_WORDS);
  uint32_t* out = tables16 + (size_t)g * COMB16_TABLE_WORDS +
                  ((size_t)w * COMB16_ENTRIES + (d0 - 1)) * COMB16_ENTRY_WORDS;
  uint32_t* sc = scratch + (size_t)t * COMB16_RUN * COMB_ENTRY_WORDS;
  G1J acc;
  pt_mul_u64_affine(acc, base, d0);
  Fp prod;
  fp_one(prod);
  for (uint32_t i = 0; i < COMB16_RUN; i++) {
    fp_mul(prod, prod, acc.z);
    uint32_t* e = out + (size_t)i * COMB16_ENTRY_WORDS;
    uint32_t* z = sc + (size_t)i * COMB_ENTRY_WORDS;
    for (int k = 0; k < FPW; k++) {
      e[k] = acc.x.v[k];
      e[FPW + k] = acc.y.v[k];
      z[k] = acc.z.v[k];
      z[FPW + k] = prod.v[k];
    }
    for (int k = 2 * FPW; k < (int)COMB16_ENTRY_WORDS; k++) e[k] = 0;
    if (i + 1 < COMB16_RUN) pt_madd(acc, acc, base);
  }
  Fp inv;
  fp_inv(inv, prod);  // 1 / (z_0 ... z_{RUN-1})
  for (int i = (int)COMB16_RUN - 1; i >= 0; i--) {
    uint32_t* e = out + (size_t)i * COMB16_ENTRY_WORDS;
    const uint32_t* z = sc + (size_t)i * COMB_ENTRY_WORDS;
    Fp zd, pre, zi;
    for (int k = 0; k < FPW; k++) zd.v[k] = z[k];
    if (i > 0) {
      for (int k = 0; k < FPW; k++) pre.v[k] = sc[(size_t)(i - 1) * COMB_ENTRY_WORDS + FPW + k];
      fp_mul(zi, inv, pre);
    } else {
      zi = inv;
    }
    fp_mul(inv, inv, zd);
    Fp X, Y, zi2, zi3;
    for (int k = 0; k < FPW; k++) {
      X.v[k] = e[k];
      Y.v[k] = e[FPW + k];
    }
    fp_sqr(zi2, zi);
    fp_mul(zi3, zi2, zi);
    G1A a;
    fp_mul(a.x, X, zi2);
    fp_mul(a.y, Y, zi3);
    store_g1a(e, a);
  }
}
void launch_comb16_build(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* tables,
                         uint32_t* tables16, uint32_t* scratch) {
  if (!count) return;
  const uint32_t n = count * COMB16_WINDOWS * (COMB16_ENTRIES / COMB16_RUN);
  hipLaunchKernelGGL(k_comb16_build, GRID(n, 64), 0, s, first, count, tables, tables16, scratch);
}

// ------------------------------------------------------------------ fixed-base comb MSM
// Each lane sums up to term_count terms of one chunk in XYZZ coordinates
// (the cheaper mixed addition) and stores one Jacobian partial.
// Scalar k (canonical, < r < 2^255) is recoded into signed digits: 16 of 16
// bits over the wide table when the generator slot is below n16, else 32 of
// 8 bits.  Windows never straddle a 32-bit word of k.  The loop is software
// pipelined: the table entry of the next (term, window) is loaded while the
// mixed addition of the current one runs, which hides the latency of the
// random reads from the wide tables.
struct CombCursor {
  const uint32_t* __restrict__ term_gen;
  const uint32_t* __restrict__ term_scalar;
  const uint32_t* __restrict__ tables;
  const uint32_t* __restrict__ tables16;
  uint32_t n16;
  uint32_t t, tend, w, carry;
  // shape of the current term
  const uint32_t* tab;
  uint32_t bits, windows, wstride, estride;
  uint32_t k[8];

  DEV void term() {
    const uint32_t g = term_gen[t];
    const bool wide = g < n16;
    bits = wide ? 16 : 8;
    windows = wide ? COMB16_WINDOWS : COMB_WINDOWS;
    wstride = wide ? COMB16_ENTRIES * (uint32_t)COMB16_ENTRY_WORDS : COMB_ENTRIES * (uint32_t)COMB_ENTRY_WORDS;
    estride = wide ? (uint32_t)COMB16_ENTRY_WORDS : (uint32_t)COMB_ENTRY_WORDS;
    tab = wide ? tables16 + (size_t)g * COMB16_TABLE_WORDS : tables + (size_t)g * COMB_TABLE_WORDS;
#pragma unroll
    for (int q = 0; q < 8; q++) k[q] = term_scalar[(size_t)t * FR_WORDS + q];
    w = 0;
    carry = 0;
  }
  // Next digit: false when the chunk is exhausted; entry = nullptr for a zero
  // digit, else the table entry of |digit| and neg its sign.
  DEV bool next(const uint4*& entry, bool& neg) {
    if (t >= tend) return false;
    const uint32_t o = w * bits;
    const uint32_t j = o >> 5;
    uint32_t word = k[0];
#pragma unroll
    for (int q = 1; q < 8; q++) word = j == (uint32_t)q ? k[q] : word;
    const uint32_t b = ((word >> (o & 31)) & ((1u << bits) - 1)) + carry;
    neg = b > (1u << (bits - 1));
    carry = neg ? 1 : 0;
    const uint32_t ad = neg ? (1u << bits) - b : b;
    entry = ad ? reinterpret_cast<const uint4*>(tab + w * wstride + (ad - 1) * estride) : nullptr;
    if (++w == windows && ++t < tend) term();
    return true;
  }
};

__global__ __launch_bounds__(128) void k_comb_msm(const CombChunk* __restrict__ chunks, uint32_t n,
                                                 const uint32_t* __restrict__ term_gen,
                                                 const uint32_t* __restrict__ term_scalar,
                                                 const uint32_t* __restrict__ tables,
                                                 const uint32_t* __restrict__ tables16, uint32_t n16,
                                                 uint32_t* __restrict__ partials,
                                                 const uint32_t* __restrict__ gate_slot,
                                                 const uint8_t* __restrict__ valid,
                                                 const uint32_t* __restrict__ list,
                                                 const uint32_t* __restrict__ count) {
  wave_priority<PRIO_MID>();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (list ? *count : n)) return;
  const uint32_t ci = list ? list[i] : i;
  if (!list && gate_slot && valid[gate_slot[ci]]) return;
  const CombChunk c = chunks[ci];
  G1Z acc;
  g1z_set_inf(acc);
  CombCursor cur;
  cur.term_gen = term_gen;
  cur.term_scalar = term_scalar;
  cur.tables = tables;
  cur.tables16 = tables16;
  cur.n16 = n16;
  cur.t = c.term_begin;
  cur.tend = c.term_begin + c.term_count;
  if (cur.t < cur.tend) cur.term();
  constexpr int Q = (int)COMB_ENTRY_WORDS / 4;
  uint4 buf[Q];
  const uint4* e4;
  bool neg;
  bool have = cur.next(e4, neg);
  if (have && e4) {
#pragma unroll
    for (int q = 0; q < Q; q++) buf[q] = e4[q];
  }
  while (have) {
    const bool add = e4 != nullptr;
    const bool sub = neg;
    G1A e;
    {
      uint32_t ew[COMB_ENTRY_WORDS];
#pragma unroll
      for (int q = 0; q < Q; q++) {
        ew[4 * q] = buf[q].x;
        ew[4 * q + 1] = buf[q].y;
        ew[4 * q + 2] = buf[q].z;
        ew[4 * q + 3] = buf[q].w;
      }
      load_g1a(e, ew);
    }
    have = cur.next(e4, neg);
    if (have && e4) {
#pragma unroll
      for (int q = 0; q < Q; q++) buf[q] = e4[q];
    }
    if (add) {
      fp_cond_neg_lazy(e.y, sub);
      g1z_madd(acc, e);
    }
  }
  G1J out;
  g1z_to_jacobian(out, acc);
  store_g1j(partials + (size_t)c.partial * G1J_WORDS, out);
}
void launch_comb_msm(hipStream_t s, const CombChunk* chunks, uint32_t n, const uint32_t* term_gen,
                     const uint32_t* term_scalar, const uint32_t* tables, const uint32_t* tables16,
                     uint32_t n16, uint32_t* partials, const uint32_t* gate_slot,
                     const uint8_t* valid, const uint32_t* list, const uint32_t* count) {
  if (!n) return;
  hipLaunchKernelGGL(k_comb_msm, GRID(n, 64), 0, s, chunks, n, term_gen, term_scalar, tables,
                     tables16, n16, partials, gate_slot, valid, list, count);
}

// ------------------------------------------------------------------ public keys
// pk_status: 0 invalid, 1 valid, 2 valid identity
__global__ __launch_bounds__(64) void k_prepare_pk(const uint32_t* __restrict__ slots, uint32_t n,
                                                  const uint8_t* __restrict__ pk_bytes,
                                                  uint8_t* __restrict__ pk_status,
                                                  uint32_t* __restrict__ pk_lines) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t slot = slots[i];
  G2J q;
  bool ok = g2_decompress(q, pk_bytes + (size_t)slot * 96);
  if (!ok) {
    pk_status[slot] = 0;
    return;
  }
  if (pt_is_inf(q)) {
    pk_status[slot] = 2;
    return;
  }
  G2A a;
  pt_to_affine(a, q);
  Line* lines = reinterpret_cast<Line*>(pk_lines + (size_t)slot * PK_LINES_WORDS);
  g2_prepare_lines(lines, a);
  pk_status[slot] = 1;
}
void launch_prepare_pk(hipStream_t s, const uint32_t* slots, uint32_t n, const uint8_t* pk_bytes,
                       uint8_t* pk_status, uint32_t* pk_lines) {
  if (!n) return;
  hipLaunchKernelGGL(k_prepare_pk, GRID(n, 64), 0, s, slots, n, pk_bytes, pk_status, pk_lines);
}

__global__ void k_prepare_neg_g2(uint32_t* lines) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  G2A q;
  g2_load_affine_const(q, bls::G2_GEN_NEG);
  g2_prepare_lines(reinterpret_cast<Line*>(lines), q);
}
void launch_prepare_neg_g2(hipStream_t s, uint32_t* lines) {
  hipLaunchKernelGGL(k_prepare_neg_g2, dim3(1), dim3(1), 0, s, lines);
}

// ------------------------------------------------------------------ test-only primitives
__global__ __launch_bounds__(64) void k_test_g1_mul(const uint8_t* pts, const uint8_t* scalars,
                                                   uint32_t n, uint8_t* out, uint8_t* ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  G1J p;
  bool good = g1_decompress(p, pts + (size_t)48 * i);
  Fr k;
  fr_from_be_bytes_raw(k, scalars + (size_t)32 * i);
  good = good && fr_canonical(k);
  ok[i] = good;
  if (!good) return;
  G1J r;
  if (pt_is_inf(p)) {
    pt_set_inf(r);
  } else {
    G1A a;
    a.x = p.x;
    a.y = p.y;
    g1_mul_glv(r, a, k.v);
  }
  uint8_t c[48];
  g1_compress(c, r);
  for (int q = 0; q < 48; q++) out[(size_t)48 * i + q] = c[q];
}
void launch_test_g1_mul(hipStream_t s, const uint8_t* pts, const uint8_t* scalars, uint32_t n,
                        uint8_t* out, uint8_t* ok) {
  if (!n) return;
  hipLaunchKernelGGL(k_test_g1_mul, GRID(n, 64), 0, s, pts, scalars, n, out, ok);
}

// e(P, Q) as a GT value through the same 12-lane Miller loop / final
// exponentiation as the pairing checks (pairing12.hpp, 5 pairs per block).
// The group's lanes share the line work: lane 0 prepares Q's lines, then
// every lane writes the line blocks (p12::line_values) of its share of the
// 68 lines x 2 pairings; the second pairing is unused (the line 1).
__global__ __launch_bounds__(64) void k_test_pairing_gt(const uint8_t* g1, const uint8_t* g2,
                                                       uint32_t n, uint8_t* out, uint8_t* ok,
                                                       uint32_t* scratch) {
  __shared__ uint32_t lds_all[p12::BLOCK_LDS_WORDS];
  const int lane = threadIdx.x;
  const int grp = lane / 12, gl = lane % 12;
  const p12::Lane L = p12::lane_setup(lds_all, lane);
  const uint32_t i = blockIdx.x * p12::CHECKS_PER_BLOCK + grp;
  const bool mine = grp < p12::CHECKS_PER_BLOCK && i < n;
  const size_t per = PK_LINES_WORDS + p12::L2_WORDS;
  uint32_t* lines = scratch + (size_t)(mine ? i : 0) * per;
  uint32_t* l2 = lines + PK_LINES_WORDS;
  bool good = false;
  G1A pa;
  fp_zero(pa.x);
  fp_zero(pa.y);
  if (mine) {
    G1J p;
    G2J q;
    good = g1_decompress(p, g1 + (size_t)48 * i) && g2_decompress(q, g2 + (size_t)96 * i);
    good = good && !pt_is_inf(p) && !pt_is_inf(q);
    if (good) {
      pa.x = p.x;
      pa.y = p.y;
      if (gl == 0) {
        G2A qa;
        pt_to_affine(qa, q);
        g2_prepare_lines(reinterpret_cast<Line*>(lines), qa);
      }
    }
  }
  __syncthreads();
  __threadfence_block();
  if (mine) {
    for (int e = gl; e < MILLER_LINES * 2; e += 12) {
      const int idx = e >> 1, q = e & 1;
      const bool use = good && q == 0;
      Fp l0, l1, c0, c1, m0, m1;
      fp_zero(m0);
      fp_zero(m1);
      fp_zero(c0);
      fp_zero(c1);
      if (use) {
        const uint32_t* line = lines + (size_t)idx * MILLER_LINE_WORDS;
        fp_load_const(l0, line);
        fp_load_const(l1, line + FPW);
        fp_load_const(c0, line + 2 * FPW);
        fp_load_const(c1, line + 3 * FPW);
        fp_mul(m0, l0, pa.x);
        fp_neg(m0, m0);
        fp_mul(m1, l1, pa.x);
        fp_neg(m1, m1);
      }
      p12::line_values(l2 + ((size_t)idx * 2 + q) * p12::LV_N * FPW, m0, m1, c0, c1, !use);
    }
  }
  __syncthreads();
  __threadfence_block();
  Fp f, r, zero;
  fp_zero(zero);
  p12::miller(f, l2, pa.y, zero, L);
  p12::final_exp(r, f, L);
  if (mine) {
    if (gl == 0) ok[i] = good;
    // slot k is tower coordinate c_{k mod 2}.c_{k div 2}; output order c0.c0, c0.c1, c0.c2, c1.c0, ..
    const int pos = (L.k & 1) * 3 + (L.k >> 1);
    Fp c;
    fp_from_mont(c, r);
    fp_to_be_bytes(out + (size_t)576 * i + 96 * pos + 48 * L.c, c);
  }
}
void launch_test_pairing_gt(hipStream_t s, const uint8_t* g1, const uint8_t* g2, uint32_t n,
                            uint8_t* out, uint8_t* ok) {
  if (!n) return;
  uint32_t* scratch = nullptr;
  const size_t per = PK_LINES_WORDS + p12::L2_WORDS;
  if (hipMalloc(&scratch, (size_t)n * per * 4) != hipSuccess) return;
  hipLaunchKernelGGL(k_test_pairing_gt,
                     dim3((n + p12::CHECKS_PER_BLOCK - 1) / p12::CHECKS_PER_BLOCK), dim3(64), 0, s,
                     g1, g2, n, out, ok, scratch);
  (void)hipStreamSynchronize(s);
  (void)hipFree(scratch);
}

__global__ __launch_bounds__(64) void k_test_g2_decompress(const uint8_t* g2, uint32_t n,
                                                          uint8_t* out, uint8_t* ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  G2J q;
  bool good = g2_decompress(q, g2 + (size_t)96 * i);
  ok[i] = good;
  if (!good) return;
  uint8_t c[96];
  g2_compress(c, q);
  for (int k = 0; k < 96; k++) out[(size_t)96 * i + k] = c[k];
}
void launch_test_g2_decompress(hipStream_t s, const uint8_t* g2, uint32_t n, uint8_t* out,
                               uint8_t* ok) {
  if (!n) return;
  hipLaunchKernelGGL(k_test_g2_decompress, GRID(n, 64), 0, s, g2, n, out, ok);
}

// Throughput probe: `iters` dependent Fp multiplications per lane.
__global__ __launch_bounds__(256) void k_fp_mul_bench(uint32_t* data, uint32_t n, uint32_t iters) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fp a, b, c, d;
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    a.v[k] = data[(size_t)i * 2 * FPW + k];
    b.v[k] = data[(size_t)i * 2 * FPW + FPW + k];
  }
  c = a;
  d = b;
  for (uint32_t t = 0; t < iters; t++) {
    fp_mul(a, a, b);
    fp_mul(c, c, d);
  }
  fp_add(a, a, c);
#pragma unroll
  for (int k = 0; k < FPW; k++) data[(size_t)i * 2 * FPW + k] = a.v[k];
}
void launch_fp_mul_bench(hipStream_t s, uint32_t* data, uint32_t n, uint32_t iters) {
  hipLaunchKernelGGL(k_fp_mul_bench, GRID(n, 256), 0, s, data, n, iters);
}

// Peak v_mad_u64_u32 throughput: 8 independent 64-bit accumulator chains per
// lane, `iters` rounds; 8 * iters multiply-adds per lane.
__global__ __launch_bounds__(256) void k_mad_peak(uint32_t* data, uint32_t n, uint32_t iters) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t x = data[i] | 1u;
  uint64_t a0 = x, a1 = x + 1, a2 = x + 2, a3 = x + 3, a4 = x + 4, a5 = x + 5, a6 = x + 6, a7 = x + 7;
  uint32_t y = x * 2654435761u;
#pragma unroll 1
  for (uint32_t t = 0; t < iters; t++) {
#pragma unroll
    for (int u = 0; u < 8; u++) {
      a0 = (uint64_t)(uint32_t)a0 * y + (a0 >> 32);
      a1 = (uint64_t)(uint32_t)a1 * y + (a1 >> 32);
      a2 = (uint64_t)(uint32_t)a2 * y + (a2 >> 32);
      a3 = (uint64_t)(uint32_t)a3 * y + (a3 >> 32);
      a4 = (uint64_t)(uint32_t)a4 * y + (a4 >> 32);
      a5 = (uint64_t)(uint32_t)a5 * y + (a5 >> 32);
      a6 = (uint64_t)(uint32_t)a6 * y + (a6 >> 32);
      a7 = (uint64_t)(uint32_t)a7 * y + (a7 >> 32);
    }
  }
  data[i] = (uint32_t)(a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7);
}
void launch_mad_peak(hipStream_t s, uint32_t* data, uint32_t n, uint32_t iters) {
  hipLaunchKernelGGL(k_mad_peak, GRID(n, 256), 0, s, data, n, iters);
}

}  // namespace cubbs
