// Descriptors shared by the host orchestration (bbs_abi.cpp) and the device
// kernels (bbs_kernels.hip), plus the launcher prototypes.  Plain structs, no
// HIP types beyond hipStream_t.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace cubbs {

// Generator numbering used everywhere on the device:
//   0 = P1 (the ciphersuite base point), 1 = Q1, 2 + j = H_{j+1}.
constexpr uint32_t GEN_P1 = 0;
constexpr uint32_t GEN_Q1 = 1;
constexpr uint32_t GEN_H0 = 2;

// Fixed-base comb: 32 windows of 8 bits, signed digits in [-128, 128],
// affine table entries d * 2^(8w) * G for d = 1..128.
constexpr int COMB_WINDOWS = 32;
constexpr int COMB_ENTRIES = 128;
constexpr size_t COMB_ENTRY_WORDS = 24;  // affine x, y (Montgomery), 96 bytes
constexpr size_t COMB_TABLE_WORDS = (size_t)COMB_WINDOWS * COMB_ENTRIES * COMB_ENTRY_WORDS;

constexpr int MILLER_LINE_WORDS = 48;   // (lambda, c) in Fp2: 192 bytes
constexpr int MILLER_LINE_COUNT = 68;
constexpr size_t PK_LINES_WORDS = (size_t)MILLER_LINE_WORDS * MILLER_LINE_COUNT;

constexpr int G1J_WORDS = 36;  // Jacobian G1 point
constexpr int G1A_WORDS = 24;  // affine G1 point
constexpr int FR_WORDS = 8;

// hash-to-scalar job: message bytes -> scalar written to out[out_index]
struct H2sJob {
  uint64_t off;
  uint32_t len;
  uint32_t out_index;
};

// domain job (crates/cubbs-bbs-protocol/src/utils/util.rs:147-195)
struct DomainJob {
  uint64_t hdr_off;
  uint32_t hdr_len;
  uint32_t pk_slot;
  uint32_t L;
  uint32_t out_index;  // scalar slot for the domain
};

// A chunk of consecutive fixed-base terms summed into one partial.
struct CombChunk {
  uint32_t term_begin;
  uint32_t term_count;
  uint32_t partial;
};

// Sum of partial slots [begin, begin + count).
struct ReduceGroup {
  uint32_t begin;
  uint32_t count;
};

// Per-item record of a batched Verify.
struct VerifyItem {
  uint64_t sig_off;      // 80-byte signature in the signature blob
  uint32_t pk_slot;
  uint32_t term_begin;   // L + 2 terms: P1, Q1, H_1..H_L
  uint32_t var_partial;  // partial slot receiving (-e) * A
  uint32_t flags;        // bit0: structurally valid (lengths)
};

// Per-item record of a batched ProofVerify.
struct ProofVerifyItem {
  uint64_t proof_off;    // proof bytes in the proof blob
  uint64_t ph_off;       // presentation header in the header blob
  uint32_t ph_len;
  uint32_t pk_slot;
  uint32_t R;            // disclosed count
  uint32_t U;            // undisclosed count
  uint32_t disc_begin;   // index of this item's first disclosed index / scalar
  uint32_t term_begin;   // T2 terms: P1, Q1, then H_1..H_L in index order
  uint32_t t2_var_partial;  // r3^ * D
  uint32_t t1_partial;      // 3 consecutive partials: c*Bbar, e^*Abar, r1^*D
  uint32_t dom_index;    // scalar slot holding the domain
  uint32_t flags;        // bit0: structurally valid
};

// Per-item record of a batched Sign.
struct SignItem {
  uint32_t term_begin;   // L + 2 terms
  uint32_t L;
  uint32_t flags;
  uint32_t pad;
  uint8_t sk[32];
};

// Per-item record of a batched ProofGen.
struct ProofGenItem {
  uint64_t sig_off;
  uint64_t ph_off;
  uint32_t ph_len;
  uint32_t pk_slot;
  uint32_t L, R, U;
  uint32_t disc_begin;     // disclosed indexes for this item
  uint32_t b_term_begin;   // L + 2 terms of B
  uint32_t t2_term_begin;  // U terms of sum H_j m~_j
  uint32_t b_group;        // reduce group index of B
  uint32_t t2_group;       // reduce group index of the H part of T2
  uint32_t rs_begin;       // 5 + U random scalars (canonical)
  uint32_t msg_scalar_begin;  // L message scalars
  uint32_t dom_index;
  uint32_t flags;
  uint64_t proof_off;      // output offset in the proof output blob
};

// ------------------------------------------------------------------ launchers
// (all asynchronous on the given stream)
void launch_h2s(hipStream_t s, const H2sJob* jobs, uint32_t n, const uint8_t* blob, int dst_kind,
                uint32_t* out_scalars);
void launch_hash_jobs(hipStream_t s, int kind, const H2sJob* jobs, uint32_t n, const uint8_t* blob,
                      const uint8_t* dst, uint32_t dst_len, uint8_t* out);
void launch_domain(hipStream_t s, const DomainJob* jobs, uint32_t n, const uint8_t* hdr_blob,
                   const uint8_t* pk_bytes, const uint8_t* gen_comp, uint32_t* out_scalars);
void launch_gen_seeds(hipStream_t s, uint32_t first, uint32_t count, uint32_t* v_state,
                      uint32_t* seeds_out);
void launch_gen_points(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* seeds,
                       uint32_t* gen_affine, uint8_t* gen_comp);
// scratch: count * COMB_WINDOWS * COMB_ENTRIES * 24 words
void launch_comb_build(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* gen_affine,
                       uint32_t* tables, uint32_t* scratch);
void launch_comb_msm(hipStream_t s, const CombChunk* chunks, uint32_t n, const uint32_t* term_gen,
                     const uint32_t* term_scalar, const uint32_t* tables, uint32_t* partials);
void launch_reduce_affine(hipStream_t s, const ReduceGroup* groups, uint32_t n,
                          const uint32_t* partials, uint32_t* out_affine, uint8_t* out_inf);
void launch_prepare_pk(hipStream_t s, const uint32_t* slots, uint32_t n, const uint8_t* pk_bytes,
                       uint8_t* pk_status, uint32_t* pk_lines);
void launch_prepare_neg_g2(hipStream_t s, uint32_t* lines);
void launch_verify_sig(hipStream_t s, const VerifyItem* items, uint32_t n, const uint8_t* sig_blob,
                       uint32_t* term_scalar, uint32_t* a_affine, uint8_t* a_inf, uint8_t* status,
                       uint32_t* partials);
void launch_verify_pairing(hipStream_t s, const VerifyItem* items, uint32_t n,
                           const uint32_t* a_affine, const uint8_t* a_inf, const uint32_t* b_affine,
                           const uint8_t* b_inf, const uint8_t* pk_status, const uint32_t* pk_lines,
                           const uint32_t* neg_g2_lines, uint8_t* status, uint8_t* valid);

// test-only primitives
void launch_test_g1_mul(hipStream_t s, const uint8_t* pts, const uint8_t* scalars, uint32_t n,
                        uint8_t* out, uint8_t* ok);
void launch_test_pairing_gt(hipStream_t s, const uint8_t* g1, const uint8_t* g2, uint32_t n,
                            uint8_t* out, uint8_t* ok);
void launch_test_g2_decompress(hipStream_t s, const uint8_t* g2, uint32_t n, uint8_t* out,
                               uint8_t* ok);
void launch_fp_mul_bench(hipStream_t s, uint32_t* data, uint32_t n, uint32_t iters);

}  // namespace cubbs
