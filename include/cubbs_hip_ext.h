/*
 * Extensions of libcubbs_bbs_hip.so beyond the reference ABI v2.
 *
 * None of these has a counterpart in the reference header
 * (native/cuda/include/cubbs_bbs_gpu_abi.h); they serve callers the reference
 * has no entry for:
 *   - reproducible ProofGen (seeded blinding, or explicitly injected scalars
 *     as in the CFRG fixtures' traces);
 *   - staged batches, so a benchmark or a serving loop can keep inputs
 *     resident on the device and launch the pipeline repeatedly;
 *   - small primitive probes used by the parity tests.
 * Same conventions as the ABI: caller-owned memory, status codes
 * CUBBS_BBS_GPU_*, abi_version must be CUBBS_BBS_GPU_ABI_VERSION.
 */
#ifndef CUBBS_HIP_EXT_H
#define CUBBS_HIP_EXT_H

#include "cubbs_bbs_gpu_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ProofGen whose blinding scalars come from ChaCha20 (RFC 8439) keyed by a
 * 32-byte seed per item, zero nonce, block counter k for scalar k; each
 * 64-byte block read as a little-endian integer and reduced mod r.  Scalar
 * order: r1, r2, e~, r1~, r3~, m~_1..m~_U (U = undisclosed count). */
int32_t cubbs_hip_proof_gen_seeded_batch(uint32_t abi_version, CuBbsBbsGpuProofGenItem *items,
                                         size_t item_count, const uint8_t *seeds32,
                                         CuBbsBbsGpuResult *results);

/* ProofGen with the blinding scalars given explicitly: item i reads
 * (5 + U_i) big-endian 32-byte scalars starting at scalars_be + 32 * offsets[i]. */
int32_t cubbs_hip_proof_gen_with_scalars_batch(uint32_t abi_version,
                                               CuBbsBbsGpuProofGenItem *items, size_t item_count,
                                               const uint8_t *scalars_be, const size_t *offsets,
                                               CuBbsBbsGpuResult *results);

/* cubbs_bbs_gpu_messages_to_scalar_batch_v2 without allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same validation, results and outputs as the allocating entry,
 * on the same hash kernel; the device memory is one static array of
 * CUBBS_HIP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends) and
 * the host staging a static buffer of the same size.  A call whose packed
 * job records, DST, message bytes and 32-byte outputs (each part rounded up
 * to 256 bytes) exceed that size returns CUBBS_BBS_GPU_ERR_INVALID_INPUT
 * without touching the device.  Calls share the array, so they are serialised
 * by a lock (overlapping calls would race on it).  Unlike the allocating entry
 * it does not record the scalars for
 * cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2. */
#define CUBBS_HIP_NOALLOC_SCRATCH_BYTES (64u << 20)
int32_t cubbs_hip_messages_to_scalar_noalloc_batch(uint32_t abi_version,
                                                   CuBbsBbsGpuMessagesToScalarItem *items,
                                                   size_t item_count, CuBbsBbsGpuResult *results);

/* ---- staged batches: prepare (host packing + upload), launch (enqueue the
 * device pipeline on the batch's own stream), sync, read results, free.
 * kind: 0 = Verify, 1 = ProofVerify, 16 = item MSMs (CuBbsBbsGpuG1MsmItem, as
 * cubbs_bbs_gpu_g1_msm_batch_v2), 17 = common-point MSM groups
 * (CuBbsBbsGpuG1MsmCommonPointsBatchItem, as
 * cubbs_bbs_gpu_g1_msm_common_points_batch_v2; point_outs are not used).
 * Items are read only during prepare.  For the MSM kinds, results hold one
 * entry per output point (per item, or per row of each group in order) and
 * cubbs_hip_batch_outputs returns the 48-byte encodings in the same order. */
#define CUBBS_HIP_BATCH_VERIFY 0
#define CUBBS_HIP_BATCH_PROOF_VERIFY 1
#define CUBBS_HIP_BATCH_G1_MSM 16
#define CUBBS_HIP_BATCH_G1_MSM_COMMON_POINTS 17
void *cubbs_hip_batch_prepare(uint32_t abi_version, int kind, const void *items, size_t item_count,
                              int32_t *status_out);
/* As cubbs_hip_batch_prepare, with flags:
 * CUBBS_HIP_BATCH_SINGLE_STREAM: the batch runs on one work stream instead of
 * forking its decode / variable-base product branch onto a second one.  A
 * single batch then takes longer, but it uses one hardware queue instead of
 * two.  Callers that keep many batches in flight get more throughput this
 * way: eight Verify+ProofVerify sets at GPU_MAX_HW_QUEUES=16 ran 8% faster
 * than four forked sets.  Bits other than the flags below are rejected
 * (code 1). */
#define CUBBS_HIP_BATCH_SINGLE_STREAM 1u
/* CUBBS_HIP_BATCH_AGGREGATE (Verify / ProofVerify): the pairing checks of
 * items under one key are also tested in groups of up to n (the group size
 * below; n + 1 when a key's last item would otherwise be checked alone) as
 * one random linear combination, e(sum c_i a_i, PK) e(sum c_i b_i, -G2) = 1, with
 * c_i = s_i + t_i lambda (s_i, t_i two 64-bit values from the OS CSPRNG drawn
 * at prepare, 2^128 distinct coefficients).  Items of a
 * group that holds are accepted without their own check; the items of a
 * failing group are checked one by one, so every per-item result equals the
 * individual check's except with probability about 2^-128 per group.  This
 * is the batch test of the reference's default same-key verify path.  It
 * lengthens one batch's chain by a pairing launch, so it pays when several
 * batches are in flight: eight single-stream sets ran 4.5 ms per step with
 * it against 6.5 ms without.  The group size is the one of CUBBS_HIP_BATCH_AGG_GROUP(n)
 * (2 <= n <= 8) when that is or-ed into the flags, else the pipeline tuning
 * table's for the batch's size class (cubbs_hip_tune_load), else 8.
 * CUBBS_HIP_BATCH_AGG_SUPER(m) (2 <= m <= 8, with CUBBS_HIP_BATCH_AGGREGATE)
 * adds a level above the groups: up to m consecutive groups of one key are
 * first checked as one combination (their c_i a_i and c_i b_i summed), and
 * their groups only when that check fails, their items only when a group
 * fails.  Without it a batch takes the table's choice for its size class
 * when it leaves the group size open too, else no supergroups.
 * CUBBS_HIP_BATCH_COMB_CHUNK(n) (1 <= n <= 64, Verify / ProofVerify only)
 * sets the fixed-base comb terms per lane of this batch (otherwise the
 * table's choice for its size class, else 4); the pipeline tuner stages its
 * candidates this way.
 * CUBBS_HIP_BATCH_AGG_STRAUS(n) (1 <= n <= 4, with CUBBS_HIP_BATCH_AGGREGATE,
 * Verify items): a group's sums sum c_i A_i and sum (-c_i e_i) A_i are formed
 * by product jobs over up to n of its signatures' A points each (one table
 * of multiples per point for both sums, one chain of doublings per sum for
 * all n points) instead of two products per item; with n >= 2 the items'
 * own products then run only for groups whose check fails (with n = 1 the
 * jobs' sums are the items' own products).  Fewer field operations per
 * item, a longer product launch.  Without it, a batch that leaves the group
 * size open too takes the table's choice for its size class, else two
 * products per item. */
#define CUBBS_HIP_BATCH_AGGREGATE 2u
#define CUBBS_HIP_BATCH_AGG_GROUP(n) (((uint32_t)(n) & 15u) << 8)
#define CUBBS_HIP_BATCH_COMB_CHUNK(n) (((uint32_t)(n) & 127u) << 12)
#define CUBBS_HIP_BATCH_AGG_SUPER(m) (((uint32_t)(m) & 15u) << 20)
#define CUBBS_HIP_BATCH_AGG_STRAUS(n) (((uint32_t)(n) & 15u) << 24)
void *cubbs_hip_batch_prepare_flags(uint32_t abi_version, int kind, const void *items,
                                    size_t item_count, uint32_t flags, int32_t *status_out);
/* A Verify batch and a ProofVerify batch staged as ONE batch: every stage
 * of the device pipeline (decode, hashing, products, comb, reductions,
 * challenge, each level of pairing checks) launches once for the items of
 * both, so each launch holds both batches' work and the pair takes one
 * stream instead of two.  Flags as cubbs_hip_batch_prepare_flags (the
 * tuning table's size class is the Verify item count).  The batch's results
 * (cubbs_hip_batch_results) are verify_count Verify results followed by
 * proof_count ProofVerify results, each as the separate batch would give. */
void *cubbs_hip_batch_prepare_verify_set(uint32_t abi_version,
                                         const CuBbsBbsGpuVerifyItem *verify_items,
                                         size_t verify_count,
                                         const CuBbsBbsGpuProofVerifyItem *proof_items,
                                         size_t proof_count, uint32_t flags, int32_t *status_out);
int32_t cubbs_hip_batch_launch(void *batch);
int32_t cubbs_hip_batch_sync(void *batch);
int32_t cubbs_hip_batch_results(void *batch, CuBbsBbsGpuResult *results);
void cubbs_hip_batch_free(void *batch);
/* on != 0: capture the batch's whole launch sequence (both streams, every
 * kernel; no memset nodes) into a HIP graph once; later launches replay it.
 * on == 0: drop the graph and launch kernel by kernel again. */
int32_t cubbs_hip_batch_use_graph(void *batch, int32_t on);
/* MSM kinds: after cubbs_hip_batch_results, copies 48 bytes per output point
 * (zeros for rows with an error) into out; *out_len = bytes written. */
int32_t cubbs_hip_batch_outputs(void *batch, uint8_t *out, size_t out_cap, size_t *out_len);
/* Kernel timing with HIP events recorded on the batch's stream around every
 * launch of `iters` back-to-back pipeline runs.  names_out receives a
 * newline-separated list of kernel names (capacity names_cap bytes); ms_out
 * receives the average milliseconds per launch of each name, in that order.
 * Returns the number of kernels (<= max_kernels), negative on error. */
int32_t cubbs_hip_batch_profile(void *batch, int iters, char *names_out, size_t names_cap,
                                double *ms_out, int max_kernels, double *total_ms_out);
/* Work units of the batch for the roofline: number of items, total terms of
 * the fixed-base MSM, mixed additions they take (16 per term over a generator
 * with wide 16-bit comb tables, 32 per term otherwise), pairings. */
int32_t cubbs_hip_batch_shape(void *batch, uint64_t *items, uint64_t *msm_terms,
                              uint64_t *comb_madds, uint64_t *pairings);

/* Algorithmic work units of a staged batch (not instruction counts), for a
 * roofline over whole pipeline steps: a cost model prices each unit in field
 * multiplications.  pair_checks_run is filled in by cubbs_hip_batch_results:
 * the group checks plus the per-item checks that were not accepted through
 * their group at that launch (items already rejected earlier are counted as
 * well, so it is an upper bound); before the first results it equals
 * group_checks + pair_checks. */
typedef struct CubbsHipBatchWork {
  uint64_t items;           /* items on the device */
  uint64_t comb_terms;      /* fixed-base MSM terms */
  uint64_t comb_madds;      /* their mixed additions: 16 per term over a wide (16-bit) table, 32 over an 8-bit table */
  uint64_t glv_products;    /* variable-base products with full scalars (two 128-bit GLV halves) */
  uint64_t agg_products;    /* variable-base products with two 64-bit halves (aggregation coefficients) */
  uint64_t decoded_points;  /* G1 decompressions (square root) with their subgroup checks */
  uint64_t point_adds;      /* Jacobian additions of the partial-sum reductions */
  uint64_t pair_checks;     /* per-item pairing checks planned (2 Miller loops + 1 final exponentiation each) */
  uint64_t group_checks;    /* random-linear-combination group checks planned */
  uint64_t pair_checks_run; /* checks computed at the last launch, see above */
  uint64_t sha_blocks;      /* SHA-256 compressions of message hashes, domains and challenges */
  uint64_t straus_sums;     /* multi-term variable-base jobs: one shared chain of 128 doublings each */
  uint64_t straus_terms;    /* their terms (full scalars): window and table additions of one GLV product each */
  uint64_t straus_short_sums;  /* second sums of dual jobs over the same points and tables, with
                                  aggregation coefficients (64-bit GLV halves): a chain of 64 doublings each */
  uint64_t straus_short_terms; /* their terms: the window additions of a 64-bit-half product each */
} CubbsHipBatchWork;
int32_t cubbs_hip_batch_work(void *batch, CubbsHipBatchWork *out);

/* Aggregated checks of a staged Verify / ProofVerify batch
 * (CUBBS_HIP_BATCH_AGGREGATE at prepare): the number of group checks it
 * holds, and how many of them held at the last cubbs_hip_batch_results.
 * Items of a group that held are accepted without their own pairing check. */
int32_t cubbs_hip_batch_group_checks(void *batch, uint64_t *groups, uint64_t *passed);

/* ---- resident state: what the library keeps on the device across calls
 * (generator points and comb tables, prepared public keys, -G2 lines, and
 * buffers replaced by a growth step that live batches may still use).  The
 * prepared keys live in a fixed table of CUBBS_HIP_PK_CACHE_ENTRIES entries
 * (default 4096; least recently used entries that no live batch uses are
 * replaced); the 8-bit comb tables are bounded by
 * CUBBS_HIP_GEN_TABLE_BUDGET_BYTES (default 32 GiB).  A batch that needs more
 * distinct keys or generator tables than these hold at once gets private
 * copies for its lifetime.  Both settings are read once, at the first call. */
typedef struct CubbsHipResidentStats {
  uint64_t resident_bytes;       /* device bytes held, retired buffers included */
  uint64_t pk_capacity;          /* entries of the key table (0 before the first key) */
  uint64_t pk_entries;           /* entries holding a key */
  uint64_t pk_pinned;            /* entries used by a live batch */
  uint64_t generator_slots;      /* generator points made (P1, Q1, H_1..) */
  uint64_t table_slots;          /* generators with a resident 8-bit comb table */
  uint64_t table_max_slots;      /* the budget, in generators */
  uint64_t wide_slots;           /* generators with a 16-bit comb table */
  uint64_t retired_buffers;      /* replaced buffers not yet freed */
  uint64_t live_batches;         /* batches holding resident state */
} CubbsHipResidentStats;
int32_t cubbs_hip_resident_stats(CubbsHipResidentStats *out);

/* ---- coalescing of small calls.  cubbs_bbs_gpu_pairing_eq_batch_v2 and
 * cubbs_bbs_gpu_g1_msm_batch_v2 calls of at most 16 items from concurrent
 * host threads are combined into one device call while earlier combined
 * calls run (at most CUBBS_HIP_COALESCE_INFLIGHT, default 2, at once); each
 * caller gets exactly its own results.  CUBBS_HIP_COALESCE=0 turns it off
 * (both read once).  Counters: combined device calls and the calls they
 * served, per entry point. */
int32_t cubbs_hip_coalesce_stats(uint64_t *pairing_combined, uint64_t *pairing_calls,
                                 uint64_t *msm_combined, uint64_t *msm_calls);

/* ---- serving loop: continuous batching of mixed requests.
 * `slots` worker slots, each with its own stream and reusable device/pinned
 * memory; a free slot takes every pending request up to max_batch and its
 * share of the HBM budget (hbm_budget_bytes / slots), runs one sub-batch per
 * operation kind and completes them, while other slots' batches run on the
 * device.  The budget covers the library's resident state too (see
 * cubbs_hip_resident_stats): a slot's share is (budget - resident bytes) /
 * slots at the time it admits a batch, and a slot always admits at least one
 * request.  A request whose estimate exceeds the whole budget is refused at
 * submission (code 1).  Zero fields take the defaults (slots: the library's
 * work streams / 8, i.e. 2 with GPU_MAX_HW_QUEUES=16; 4096; the resident
 * bytes at creation + 4 GiB). */
typedef struct CubbsHipServerConfig {
  uint32_t slots;
  uint32_t max_batch;
  uint64_t hbm_budget_bytes;
  uint32_t flags;  /* CUBBS_HIP_SERVER_* below */
  /* The request batcher of the reference's W3C service
   * (crates/cubbs-w3c-service/src/batcher.rs:22-36, 222-300), in the loop:
   * coalesce_us is the window a free slot waits for a partial batch (fewer
   * than max_batch pending) to fill before it takes it, one slot collecting
   * at a time (0: take what is pending at once); queue_capacity bounds the
   * pending requests, and a submit blocks while the queue is full (0:
   * unbounded). */
  uint32_t coalesce_us;
  uint32_t queue_capacity;
  uint32_t reserved;
} CubbsHipServerConfig;
/* Each sub-batch's launch sequence is captured into a HIP graph, and the
 * slot's instantiated graph of that kind is updated from it
 * (hipGraphExecUpdate) and launched, instead of launching kernel by kernel;
 * a new graph is instantiated when the sequence's shape changes. */
#define CUBBS_HIP_SERVER_GRAPHS 1u
/* Verify / ProofVerify sub-batches check their pairings in per-key groups of
 * 8 (as CUBBS_HIP_BATCH_AGGREGATE). */
#define CUBBS_HIP_SERVER_AGGREGATE 2u

#define CUBBS_HIP_REQ_VERIFY 0
#define CUBBS_HIP_REQ_PROOF_VERIFY 1
#define CUBBS_HIP_REQ_SIGN 2
#define CUBBS_HIP_REQ_PROOF_GEN 3

/* One request: `item` points to the ABI item of the kind (VerifyItem,
 * ProofVerifyItem, SignItem, ProofGenItem); output buffers inside the item are
 * ignored (the server keeps the output until it is waited for).  ProofGen
 * blinding: ChaCha20 with seed32 as in cubbs_hip_proof_gen_seeded_batch, or
 * the OS CSPRNG when seed32 is NULL. */
typedef struct CubbsHipRequest {
  int32_t kind;
  const void *item;
  const uint8_t *seed32;
} CubbsHipRequest;

void *cubbs_hip_server_create(const CubbsHipServerConfig *cfg, int32_t *status_out);
/* Copies the request and queues it.  Returns a ticket >= 0, or -code. */
int64_t cubbs_hip_server_submit(void *server, const CubbsHipRequest *request);
/* Blocks until the ticket completes; writes its result, its output bytes
 * (signature or proof, result->bytes_written of them) and the latency from
 * submission to completion.  A ticket can be waited for once. */
int32_t cubbs_hip_server_wait(void *server, int64_t ticket, CuBbsBbsGpuResult *result,
                              uint8_t *out, size_t out_cap, uint64_t *latency_ns);
/* Open-loop load: submits request i at arrival_ns[i] (non-decreasing, from the
 * call's start) on the calling thread, then waits for all.  Latency is measured
 * from the scheduled arrival.  outs/out_caps may be NULL (outputs dropped). */
int32_t cubbs_hip_server_run_open_loop(void *server, const CubbsHipRequest *requests,
                                       const uint64_t *arrival_ns, size_t n,
                                       CuBbsBbsGpuResult *results, uint64_t *latency_ns,
                                       uint8_t *const *outs, const size_t *out_caps);
/* Counters since creation; returns the server's fault code (0 = healthy). */
int32_t cubbs_hip_server_stats(void *server, uint64_t *batches, uint64_t *requests,
                               uint64_t *max_concurrent_batches, uint64_t *peak_slot_bytes);
/* Averages per batch since creation: host time planning and packing the
 * sub-batches, and the whole batch from its first plan to its last result. */
int32_t cubbs_hip_server_timing(void *server, double *plan_ms_per_batch, double *batch_ms_per_batch);
/* Batcher statistics since creation (the reference batcher's snapshot,
 * crates/cubbs-w3c-service/src/batcher.rs:195-220): requests queued now, the
 * summed wait of every dispatched request from its submission to its batch's
 * dispatch, the largest and the last batch, and hist[k] = batches of k
 * requests for k < hist_len (larger batches counted in hist[hist_len - 1]);
 * hist may be NULL.  Returns the server's fault code (0 = healthy). */
int32_t cubbs_hip_server_queue_stats(void *server, uint64_t *queued, uint64_t *total_wait_ns,
                                     uint64_t *max_batch, uint64_t *last_batch, uint64_t *hist,
                                     size_t hist_len);
/* Drains queued requests, joins the workers and frees the server. */
void cubbs_hip_server_destroy(void *server);

/* ---- launch-shape autotuner.  Times candidate shapes on synthetic inputs on
 * the current device and keeps the fastest per size class for this process:
 * Pippenger window width and entries per piece for MSMs of 2^10 ..
 * 2^max_log2_terms terms (10 <= max_log2_terms <= 20), comb terms per lane
 * for batches of 2^12 .. 2^18 fixed-base terms.  Writes one line per class
 * ("pippenger log2_terms=K c=C S=S ms=T" / "comb log2_terms=K chunk=N ms=T")
 * into report (NUL-terminated, truncated to report_cap). */
int32_t cubbs_hip_autotune(uint32_t max_log2_terms, char *report, size_t report_cap);
/* Installs a table in the report format (e.g. saved from an earlier run);
 * an empty string clears the table.  A malformed line fails the call.  The
 * table also takes whole-operation pipeline lines,
 * "pipeline kind=verify|proof_verify log2_items=K chunk=N group=G [super=M]
 * [straus=S]", measured by the caller with sets of such batches in flight:
 * batches of about 2^K items then plan their comb with N terms per lane and,
 * when aggregated without an explicit group, groups of G (2 <= G <= 8),
 * supergroups of M groups (0: none, else 2 <= M <= 8) and, for Verify items,
 * the groups' sums in product jobs of S signatures (as
 * CUBBS_HIP_BATCH_AGG_STRAUS(S); 0: two products per item). */
int32_t cubbs_hip_tune_load(const char *table);

/* Launch shapes forced on the calling thread for the calls it makes next
 * (parity tests of every Pippenger window width and comb chunk): pip_c
 * Pippenger window width (4..16), pip_s entries per piece (1..256), comb_chunk
 * comb terms per lane (1..64); 0 = not forced.  Other threads are not
 * affected. */
int32_t cubbs_hip_test_force_shapes(uint32_t pip_c, uint32_t pip_s, uint32_t comb_chunk);

/* ---- primitive probes (parity tests) */
/* out = k * P for n compressed points and 32-byte big-endian scalars; ok[i] = 0
 * when the point or scalar does not decode. */
int32_t cubbs_hip_test_g1_mul(const uint8_t *points48, const uint8_t *scalars32, size_t n,
                              uint8_t *out48, uint8_t *ok);
/* e(P, Q) after final exponentiation (the pairing cubed, see pairing.hpp),
 * 12 Fp coefficients of 48 bytes, order c0.c0.c0, c0.c0.c1, c0.c1.c0 ... */
int32_t cubbs_hip_test_pairing_gt(const uint8_t *g1_48, const uint8_t *g2_96, size_t n,
                                  uint8_t *out576, uint8_t *ok);
/* decompress + subgroup check + recompress of G2 points */
int32_t cubbs_hip_test_g2_roundtrip(const uint8_t *g2_96, size_t n, uint8_t *out96, uint8_t *ok);
/* compressed generators [P1, Q1, H_1 .. H_{count-2}] as built on the device */
int32_t cubbs_hip_generators(size_t count, uint8_t *out48);
/* PK = sk * G2 for n 32-byte big-endian secret keys (keys.rs:256-260 of the
 * reference protocol crate); ok[i] = 0 for a non-canonical key. */
int32_t cubbs_hip_sk_to_pk(const uint8_t *sk32, size_t n, uint8_t *pk96, uint8_t *ok);
/* Fp multiplication throughput probe: lanes x iters x 2 dependent chains;
 * returns elapsed milliseconds of the kernel (HIP events) in *ms. */
int32_t cubbs_hip_fp_mul_bench(uint32_t lanes, uint32_t iters, double *ms);
/* Peak v_mad_u64_u32 rate probe: lanes x iters x 64 independent multiply-adds;
 * elapsed kernel milliseconds in *ms. */
int32_t cubbs_hip_mad_peak_bench(uint32_t lanes, uint32_t iters, double *ms);

#ifdef __cplusplus
}
#endif
#endif
