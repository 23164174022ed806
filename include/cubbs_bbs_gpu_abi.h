/*
 * C ABI v2 of the BBS GPU backend, as exported by libcubbs_bbs_hip.so.
 *
 * This is the drop-in boundary: the struct layouts, constants and entry points
 * are the ones the cuBBS-Cred Rust crates bind (crates/cubbs-bbs/src/gpu_abi.rs,
 * loaded through CUBBS_BBS_GPU_LIBRARY by crates/cubbs-bbs/src/gpu_runtime.rs:269-453).
 * Every declaration cites the reference header line it must stay binary-
 * compatible with (native/cuda/include/cubbs_bbs_gpu_abi.h).  Layouts are LP64.
 *
 * Calling convention of every batch entry point:
 *   int32_t fn(uint32_t abi_version, Item *items, size_t n, CuBbsBbsGpuResult *results)
 *   - abi_version != 2, or a NULL array with n > 0      -> CUBBS_BBS_GPU_ERR_INVALID_INPUT
 *   - n == 0                                            -> CUBBS_BBS_GPU_SUCCESS
 *   - structural errors (wrong fixed lengths, small output capacity in
 *     primitive entries)                                -> call-level error code
 *   - per-item crypto / encoding errors                 -> results[i].code
 *   - device unavailable or a device failure            -> CUBBS_BBS_GPU_ERR_INTERNAL
 * The caller owns all memory; the library keeps only content-keyed copies
 * (generator tables, prepared public keys) and never retains caller pointers.
 * Entry points are re-entrant: each calling thread gets its own HIP stream.
 */
#ifndef CUBBS_BBS_GPU_ABI_H
#define CUBBS_BBS_GPU_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference header :11 */
#define CUBBS_BBS_GPU_ABI_VERSION 2u

/* MSM route ids reported in CuBbsBbsGpuResult.value (reference :16-23) */
#define CUBBS_BBS_MSM_ROUTE_CUBBS_ITEM 1u
#define CUBBS_BBS_MSM_ROUTE_CUBBS_COMMON_POINTS 2u
#define CUBBS_BBS_MSM_ROUTE_CUBBS_LARGE 3u
#define CUBBS_BBS_MSM_ROUTE_CUBBS_UNIVERSAL 4u
#define CUBBS_BBS_MSM_ROUTE_BLST_ITEM 5u
#define CUBBS_BBS_MSM_ROUTE_BLST_COMMON_POINTS 6u
#define CUBBS_BBS_MSM_ROUTE_SPPARK_ITEM_REFERENCE 7u
#define CUBBS_BBS_MSM_ROUTE_SPPARK_COMMON_POINTS_REFERENCE 8u

/* status codes (reference :25-28) */
#define CUBBS_BBS_GPU_SUCCESS 0
#define CUBBS_BBS_GPU_ERR_INVALID_INPUT 1
#define CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT 2
#define CUBBS_BBS_GPU_ERR_INTERNAL 3

/* backend kinds (reference :30-32).  This library reports kind 1: the Rust
 * loader gates its accelerated paths on that value (gpu_runtime.rs:483-505). */
#define CUBBS_BBS_GPU_BACKEND_KIND_UNKNOWN 0u
#define CUBBS_BBS_GPU_BACKEND_KIND_CUDA 1u
#define CUBBS_BBS_GPU_BACKEND_KIND_CPU_ABI_TEST 2u

/* fixed lengths (reference :34-37) */
#define CUBBS_BBS_SECRET_KEY_LEN 32u
#define CUBBS_BBS_PUBLIC_KEY_LEN 96u
#define CUBBS_BBS_SIGNATURE_LEN 80u
#define CUBBS_BBS_2023_HEADER_LEN 64u

/* reference :39-42 */
typedef struct CuBbsBbsGpuByteSlice {
  const uint8_t *data;
  size_t len;
} CuBbsBbsGpuByteSlice;

/* reference :44-48 */
typedef struct CuBbsBbsGpuByteBuffer {
  uint8_t *data;
  size_t capacity;
  size_t len;
} CuBbsBbsGpuByteBuffer;

/* reference :50-53 */
typedef struct CuBbsBbsGpuIndexSlice {
  const size_t *data;
  size_t len;
} CuBbsBbsGpuIndexSlice;

/* reference :55-59 */
typedef struct CuBbsBbsGpuResult {
  int32_t code;
  uint8_t value;
  size_t bytes_written;
} CuBbsBbsGpuResult;

/* reference :61-64 */
typedef struct CuBbsBbsGpuBackendInfo {
  uint32_t backend_kind;
  uint8_t hardware_acceleration;
} CuBbsBbsGpuBackendInfo;

/* reference :66-73 */
typedef struct CuBbsBbsGpuSignItem {
  CuBbsBbsGpuByteSlice secret_key;
  CuBbsBbsGpuByteSlice public_key;
  CuBbsBbsGpuByteSlice header;
  const CuBbsBbsGpuByteSlice *messages;
  size_t message_count;
  CuBbsBbsGpuByteBuffer signature_out;
} CuBbsBbsGpuSignItem;

/* reference :75-81 */
typedef struct CuBbsBbsGpuVerifyItem {
  CuBbsBbsGpuByteSlice public_key;
  CuBbsBbsGpuByteSlice signature;
  CuBbsBbsGpuByteSlice header;
  const CuBbsBbsGpuByteSlice *messages;
  size_t message_count;
} CuBbsBbsGpuVerifyItem;

/* reference :83-92 */
typedef struct CuBbsBbsGpuProofGenItem {
  CuBbsBbsGpuByteSlice public_key;
  CuBbsBbsGpuByteSlice signature;
  CuBbsBbsGpuByteSlice header;
  CuBbsBbsGpuByteSlice presentation_header;
  const CuBbsBbsGpuByteSlice *messages;
  size_t message_count;
  CuBbsBbsGpuIndexSlice disclosed_indexes;
  CuBbsBbsGpuByteBuffer proof_out;
} CuBbsBbsGpuProofGenItem;

/* reference :94-103 */
typedef struct CuBbsBbsGpuProofVerifyItem {
  CuBbsBbsGpuByteSlice public_key;
  CuBbsBbsGpuByteSlice proof;
  CuBbsBbsGpuByteSlice header;
  CuBbsBbsGpuByteSlice presentation_header;
  const CuBbsBbsGpuByteSlice *disclosed_messages;
  size_t disclosed_count;
  CuBbsBbsGpuIndexSlice disclosed_indexes;
  size_t total_message_count;
} CuBbsBbsGpuProofVerifyItem;

/* reference :105-109 */
typedef struct CuBbsBbsGpuMessagesToScalarItem {
  CuBbsBbsGpuByteSlice message;
  CuBbsBbsGpuByteSlice dst;
  CuBbsBbsGpuByteBuffer scalar_out;
} CuBbsBbsGpuMessagesToScalarItem;

/* reference :111-119 */
typedef struct CuBbsBbsGpuSignEFromResidentScalarsItem {
  CuBbsBbsGpuByteSlice secret_key;
  CuBbsBbsGpuByteSlice domain;
  const CuBbsBbsGpuByteSlice *messages;
  size_t message_scalar_offset;
  size_t message_count;
  CuBbsBbsGpuByteBuffer message_scalars_out;
  CuBbsBbsGpuByteBuffer scalar_out;
} CuBbsBbsGpuSignEFromResidentScalarsItem;

/* reference :121-127 */
typedef struct CuBbsBbsGpuPrepareBItem {
  CuBbsBbsGpuByteSlice public_key;
  CuBbsBbsGpuByteSlice header;
  const CuBbsBbsGpuByteSlice *message_scalars;
  size_t message_count;
  CuBbsBbsGpuByteBuffer b_out;
} CuBbsBbsGpuPrepareBItem;

/* reference :129-134 */
typedef struct CuBbsBbsGpuG1MsmItem {
  const CuBbsBbsGpuByteSlice *points;
  const CuBbsBbsGpuByteSlice *scalars;
  size_t term_count;
  CuBbsBbsGpuByteBuffer point_out;
} CuBbsBbsGpuG1MsmItem;

/* reference :136-144: one group = one point set, output_count scalar rows
 * given either as slices (scalar_rows, term_count slices per row) or as one
 * packed big-endian array (scalar_rows_be, row stride scalar_row_stride). */
typedef struct CuBbsBbsGpuG1MsmCommonPointsBatchItem {
  const CuBbsBbsGpuByteSlice *points;
  size_t term_count;
  const CuBbsBbsGpuByteSlice *scalar_rows;
  const uint8_t *scalar_rows_be;
  size_t scalar_row_stride;
  size_t output_count;
  CuBbsBbsGpuByteBuffer *point_outs;
} CuBbsBbsGpuG1MsmCommonPointsBatchItem;

/* reference :146-154 */
typedef struct CuBbsBbsGpuG1MsmCommonPointsPlusVariableItem {
  const CuBbsBbsGpuByteSlice *points;
  size_t term_count;
  const CuBbsBbsGpuByteSlice *scalar_rows;
  const CuBbsBbsGpuByteSlice *variable_points;
  const CuBbsBbsGpuByteSlice *variable_scalars;
  size_t output_count;
  CuBbsBbsGpuByteBuffer *point_outs;
} CuBbsBbsGpuG1MsmCommonPointsPlusVariableItem;

/* reference :156-170:
 *   T1 = challenge*Bbar + e_cap*Abar + r1_cap*D
 *   T2 = MSM(common_points, common_scalars) + r3_cap*D
 * proof_points row-major [Bbar, Abar, D]; proof_scalars row-major
 * [challenge, e_cap, r1_cap, r3_cap]. */
typedef struct CuBbsBbsGpuProofVerifyMsmPairItem {
  const CuBbsBbsGpuByteSlice *common_points;
  size_t common_term_count;
  const CuBbsBbsGpuByteSlice *common_scalar_rows;
  const CuBbsBbsGpuByteSlice *proof_points;
  const CuBbsBbsGpuByteSlice *proof_scalars;
  size_t output_count;
  CuBbsBbsGpuByteBuffer *t1_outs;
  CuBbsBbsGpuByteBuffer *t2_outs;
} CuBbsBbsGpuProofVerifyMsmPairItem;

/* reference :172-179 */
typedef struct CuBbsBbsGpuG1DualMsmSamePointsItem {
  const CuBbsBbsGpuByteSlice *points;
  const CuBbsBbsGpuByteSlice *lhs_scalars;
  const CuBbsBbsGpuByteSlice *rhs_scalars;
  size_t term_count;
  CuBbsBbsGpuByteBuffer lhs_out;
  CuBbsBbsGpuByteBuffer rhs_out;
} CuBbsBbsGpuG1DualMsmSamePointsItem;

/* reference :181-189 */
typedef struct CuBbsBbsGpuG1DualMsmSamePointsLanesItem {
  const CuBbsBbsGpuByteSlice *points;
  const CuBbsBbsGpuByteSlice *lhs_scalars;
  const CuBbsBbsGpuByteSlice *rhs_scalars;
  size_t term_count;
  size_t lane_count;
  CuBbsBbsGpuByteBuffer lhs_outs;
  CuBbsBbsGpuByteBuffer rhs_outs;
} CuBbsBbsGpuG1DualMsmSamePointsLanesItem;

/* reference :191-195 */
typedef struct CuBbsBbsGpuG1AddItem {
  CuBbsBbsGpuByteSlice lhs;
  CuBbsBbsGpuByteSlice rhs;
  CuBbsBbsGpuByteBuffer point_out;
} CuBbsBbsGpuG1AddItem;

/* reference :197-202: checks e(lhs_g1, lhs_g2) == e(rhs_g1, rhs_g2) */
typedef struct CuBbsBbsGpuPairingEqItem {
  CuBbsBbsGpuByteSlice lhs_g1;
  CuBbsBbsGpuByteSlice lhs_g2;
  CuBbsBbsGpuByteSlice rhs_g1;
  CuBbsBbsGpuByteSlice rhs_g2;
} CuBbsBbsGpuPairingEqItem;

/* reference :204-210 */
typedef struct CuBbsBbsGpuPairingEqScaledItem {
  CuBbsBbsGpuByteSlice lhs_g1;
  CuBbsBbsGpuByteSlice lhs_g2;
  CuBbsBbsGpuByteSlice rhs_g1;
  CuBbsBbsGpuByteSlice rhs_g2;
  CuBbsBbsGpuByteSlice coefficient;
} CuBbsBbsGpuPairingEqScaledItem;

/* reference :212-225: d = B*r2, abar = A*(r1 r2), bbar = d*r1 - abar*e,
 * t1 = abar*e_tilde + d*r1_tilde; scalars_be = [r2, r1r2, r1r2, r1r2e,
 * r1r2e_tilde, r2r1_tilde, r3_tilde]; seed reserved. */
typedef struct CuBbsBbsGpuProofGenT2PrepareItem {
  CuBbsBbsGpuByteSlice compressed_b;
  CuBbsBbsGpuByteSlice signature_a;
  CuBbsBbsGpuByteSlice scalars_be[7];
  CuBbsBbsGpuByteSlice seed;
  CuBbsBbsGpuByteBuffer d_out;
  CuBbsBbsGpuByteBuffer abar_out;
  CuBbsBbsGpuByteBuffer bbar_out;
  CuBbsBbsGpuByteBuffer t1_out;
} CuBbsBbsGpuProofGenT2PrepareItem;

/* ------------------------------------------------------------------------
 * Entry points (reference :227-308).  The four whole-operation entries are
 * the batched device pipelines of this library (the reference leaves them as
 * stubs, native/cuda/src/cubbs_bbs_cuda.cu:826-835,1169-1200).
 * ------------------------------------------------------------------------ */
int32_t cubbs_bbs_gpu_backend_info_v2(uint32_t abi_version, CuBbsBbsGpuBackendInfo *info_out);

/* Sign: results[i].code per item; signature_out gets A || e (80 bytes). */
int32_t cubbs_bbs_gpu_sign_batch_v2(uint32_t abi_version, CuBbsBbsGpuSignItem *items,
                                    size_t item_count, CuBbsBbsGpuResult *results);

/* Verify: results[i].value = 1 valid / 0 invalid when results[i].code == 0. */
int32_t cubbs_bbs_gpu_verify_batch_v2(uint32_t abi_version, const CuBbsBbsGpuVerifyItem *items,
                                      size_t item_count, CuBbsBbsGpuResult *results);

/* ProofGen with blinding scalars from the operating system's CSPRNG. */
int32_t cubbs_bbs_gpu_proof_gen_batch_v2(uint32_t abi_version, CuBbsBbsGpuProofGenItem *items,
                                         size_t item_count, CuBbsBbsGpuResult *results);

int32_t cubbs_bbs_gpu_proof_verify_batch_v2(uint32_t abi_version,
                                            const CuBbsBbsGpuProofVerifyItem *items,
                                            size_t item_count, CuBbsBbsGpuResult *results);

/* reference :250-252; all items must share one DST (reference
 * cubbs_bbs_cuda.cu:400-410). */
int32_t cubbs_bbs_gpu_messages_to_scalar_batch_v2(uint32_t abi_version,
                                                  CuBbsBbsGpuMessagesToScalarItem *items,
                                                  size_t item_count, CuBbsBbsGpuResult *results);

/* Test exports of the reference's native smoke tests (cubbs_bbs_cuda.cu:53,117,230):
 * packed messages (offsets/lengths), one DST. */
int32_t cubbs_bbs_gpu_sha256_batch_for_test_v2(uint32_t abi_version, const uint8_t *packed,
                                               const size_t *offsets, const size_t *lengths,
                                               size_t item_count, uint8_t *digests_out);
int32_t cubbs_bbs_gpu_expand_xmd_sha256_48_batch_for_test_v2(
    uint32_t abi_version, const uint8_t *packed, const size_t *offsets, const size_t *lengths,
    size_t item_count, const uint8_t *dst, size_t dst_len, uint8_t *out48);
int32_t cubbs_bbs_gpu_hash_to_scalar_sha256_batch_for_test_v2(
    uint32_t abi_version, const uint8_t *packed, const size_t *offsets, const size_t *lengths,
    size_t item_count, const uint8_t *dst, size_t dst_len, uint8_t *scalars_out);

/* Primitive entry points (reference :254-308).  Uniform error convention:
 * structural errors (lengths, NULL pointers, output capacity) fail the call
 * with CUBBS_BBS_GPU_ERR_INVALID_INPUT; an undecodable or non-subgroup point
 * and a non-canonical scalar give results[row].code = 1; results[row].value
 * is the MSM route id (1 item, 2 common points, 3 large, 6 dual) or the
 * boolean of a pairing equality. */

/* e = hash_to_scalar(sk || m_1 .. m_L || domain, e_dst) per item, the message
 * scalars hashed from `messages` with message_dst (raw form, all items) or
 * taken from this thread's last hash_to_scalar / messages_to_scalar call
 * starting at message_scalar_offset (resident form, `messages` NULL). */
/* reference :254-257 */
int32_t cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuSignEFromResidentScalarsItem *items, size_t item_count,
    CuBbsBbsGpuByteSlice message_dst, CuBbsBbsGpuByteSlice e_dst, CuBbsBbsGpuResult *results);

/* reference :259-262 */
int32_t cubbs_bbs_gpu_prepare_b_batch_v2(uint32_t abi_version, CuBbsBbsGpuPrepareBItem *items,
                                         size_t item_count, CuBbsBbsGpuResult *results);

/* reference :264-267 */
int32_t cubbs_bbs_gpu_g1_msm_batch_v2(uint32_t abi_version, CuBbsBbsGpuG1MsmItem *items,
                                      size_t item_count, CuBbsBbsGpuResult *results);

/* reference :269-275 */
int32_t cubbs_bbs_gpu_g1_msm_common_points_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1MsmCommonPointsBatchItem *items, size_t group_count,
    CuBbsBbsGpuResult *results);
int32_t cubbs_bbs_gpu_g1_msm_fused_common_points_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1MsmCommonPointsBatchItem *items, size_t group_count,
    CuBbsBbsGpuResult *results);

/* reference :277-280 */
int32_t cubbs_bbs_gpu_g1_msm_common_points_plus_variable_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1MsmCommonPointsPlusVariableItem *items, size_t group_count,
    CuBbsBbsGpuResult *results);

/* reference :282-284 */
int32_t cubbs_bbs_gpu_proof_verify_msm_pair_batch_v2(uint32_t abi_version,
                                                     CuBbsBbsGpuProofVerifyMsmPairItem *items,
                                                     size_t group_count,
                                                     CuBbsBbsGpuResult *results);

/* reference :286-292 */
int32_t cubbs_bbs_gpu_g1_dual_msm_same_points_batch_v2(uint32_t abi_version,
                                                       CuBbsBbsGpuG1DualMsmSamePointsItem *items,
                                                       size_t item_count,
                                                       CuBbsBbsGpuResult *results);
int32_t cubbs_bbs_gpu_g1_dual_msm_same_points_lanes_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1DualMsmSamePointsLanesItem *items, size_t item_count,
    CuBbsBbsGpuResult *results);

/* reference :294-297 */
int32_t cubbs_bbs_gpu_g1_add_batch_v2(uint32_t abi_version, CuBbsBbsGpuG1AddItem *items,
                                      size_t item_count, CuBbsBbsGpuResult *results);

/* reference :299-301 */
int32_t cubbs_bbs_gpu_pairing_eq_batch_v2(uint32_t abi_version,
                                          const CuBbsBbsGpuPairingEqItem *items,
                                          size_t item_count, CuBbsBbsGpuResult *results);

/* reference :302-304: writes ONE result for the whole aggregate. */
int32_t cubbs_bbs_gpu_pairing_eq_aggregate_v2(uint32_t abi_version,
                                              const CuBbsBbsGpuPairingEqScaledItem *items,
                                              size_t item_count, CuBbsBbsGpuResult *result);

/* reference :306-308 */
int32_t cubbs_bbs_gpu_proof_gen_t2_prepare_batch_v2(uint32_t abi_version,
                                                    CuBbsBbsGpuProofGenT2PrepareItem *items,
                                                    size_t item_count,
                                                    CuBbsBbsGpuResult *results);


#ifdef __cplusplus
}
#endif

#endif
