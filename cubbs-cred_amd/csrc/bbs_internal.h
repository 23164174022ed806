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
constexpr size_t COMB_ENTRY_WORDS = 28;  // affine x, y: 2 x 14 limbs of 28 bits (Montgomery), 112 bytes
constexpr size_t COMB_TABLE_WORDS = (size_t)COMB_WINDOWS * COMB_ENTRIES * COMB_ENTRY_WORDS;

// Wide fixed-base comb for the first generator slots: 16 windows of 16 bits,
// signed digits in [-2^15, 2^15], affine entries d * 2^(16w) * G for
// d = 1..2^15, each padded to 32 words so that an entry is one aligned
// 128-byte line (64 MiB per generator).  Half the mixed additions of the
// 8-bit comb per term; the table is read at one random line per window.
constexpr int COMB16_WINDOWS = 16;
constexpr int COMB16_ENTRIES = 32768;
constexpr size_t COMB16_ENTRY_WORDS = 32;
constexpr size_t COMB16_TABLE_WORDS = (size_t)COMB16_WINDOWS * COMB16_ENTRIES * COMB16_ENTRY_WORDS;
constexpr uint32_t COMB16_RUN = 256;             // consecutive entries per lane of the build
constexpr uint32_t COMB16_DEFAULT_SLOTS = 130;   // P1, Q1, H_1..H_128

constexpr int MILLER_LINE_WORDS = 56;   // (lambda, c) in Fp2: 4 x 14 limbs, 224 bytes
constexpr int MILLER_LINE_COUNT = 68;
constexpr size_t PK_LINES_WORDS = (size_t)MILLER_LINE_WORDS * MILLER_LINE_COUNT;

constexpr int G1J_WORDS = 42;  // Jacobian G1 point, 3 x 14 limbs
constexpr int G1A_WORDS = 28;  // affine G1 point, 2 x 14 limbs
constexpr int FR_WORDS = 8;

// hash-to-scalar job: message bytes -> scalar written to out[out_index]
struct H2sJob {
  uint64_t off;
  uint32_t len;
  uint32_t out_index;
};

// Message hash of a batch pipeline (hash_to_scalar with the MAP DST): blob +
// off holds the input of expand_message_xmd's b_0 after its Z_pad block --
// msg || I2OSP(48, 2) || 0x00 || DST' and the SHA-256 padding -- as `blocks`
// 64-byte blocks of big-endian 32-bit words, 16-byte aligned; the scalar goes
// to out[out_index].
struct MsgHashJob {
  uint64_t off;
  uint32_t blocks;
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

// Compressed 48-byte encoding of points[point] (identity aware) written at
// out + off.
struct CompressJob {
  uint32_t point;
  uint32_t pad;
  uint64_t off;
};

// Decode a compressed G1 point from the input blob into points[out]; a
// failure marks status[item] = 1 (invalid input).
struct DecodeG1Job {
  uint64_t off;
  uint32_t out;
  uint32_t item;
};

// Decode a 32-byte big-endian scalar (canonical check) into scalars[out];
// flags bit0: store r - s instead of s; bit1: reduce mod r instead of
// rejecting a non-canonical value; bit2: multiply by the canonical 32-byte
// big-endian scalar at blob + mul_off (before the negation of bit0).
struct DecodeFrJob {
  uint64_t off;
  uint32_t out;
  uint32_t item;
  uint32_t flags;
  uint32_t pad;
  uint64_t mul_off;
};

// partials[partial] = scalars[scalar] * points[point] (identity-aware), on
// one lane: the GLV split k = k1 + k2 lambda evaluated as k1 P + k2 phi(P)
// with shared doublings and signed 4-bit windows (Straus), over two tables of
// affine multiples 1..8 built in a per-job workspace of VAR_WS_WORDS words.
constexpr uint32_t VAR_PARTIALS = 1;
constexpr uint32_t VAR_TABLE_WORDS = 8 * 28;                      // 8 affine points
constexpr uint32_t VAR_WS_WORDS = 2 * VAR_TABLE_WORDS + 7 * 42;   // + 7 Jacobian temps
struct VarJob {
  uint32_t point;
  uint32_t scalar;
  uint32_t partial;
};
// partials[partial] = the sum over t < count of scalars[scalar[t]] *
// points[point[t]] (identity-aware) on one lane: the GLV halves of all terms
// share one chain of doublings (Straus over 2 * count half-points).  Its
// workspace (tables of every term and the temporaries) is
// var_sum_ws_words(count) words from word `ws` of the sums' workspace.
// VAR_SUM_DUAL: the job forms a second sum over the same points and tables,
// partials[partial2] = sum of scalars[scalar2[t]] * points[point[t]] (its
// chain of doublings starts at the highest non-zero window of those scalars,
// so short scalars, such as aggregation coefficients, take a short chain).
// VAR_SUM_GATED: the job runs only when valid[gate] == 0 (its group's check
// did not hold); otherwise it writes nothing.
constexpr uint32_t VAR_TERMS_MAX = 4;
constexpr uint32_t VAR_SUM_DUAL = 1, VAR_SUM_GATED = 2;
struct VarSum {
  uint32_t point[VAR_TERMS_MAX];
  uint32_t scalar[VAR_TERMS_MAX];
  uint32_t scalar2[VAR_TERMS_MAX];
  uint32_t count;
  uint32_t partial;
  uint32_t partial2;
  uint32_t ws;
  uint32_t flags;
  uint32_t gate;
  uint32_t pad[2];
};
constexpr uint32_t var_sum_ws_words(uint32_t count) { return count * 2 * VAR_TABLE_WORDS + 7 * 42; }

// Fill one fixed-base term scalar: mode 0: scalars[dst] = scalars[a];
// mode 1: scalars[dst] = scalars[a] * scalars[b] mod r; mode 2: scalars[dst] = 1;
// mode 3: scalars[dst] = sum over k in [a, a + b) of scalars[x_k] * scalars[y_k]
// mod r, where (x_k, y_k) = (sum_pairs[2k], sum_pairs[2k + 1]) and y_k =
// TERM_SUM_ONE stands for the factor 1.
constexpr uint32_t TERM_SUM_ONE = 0xffffffffu;
struct TermFill {
  uint32_t dst;
  uint32_t a;
  uint32_t b;
  uint32_t mode;
};

// points[out] = partials[extra] (unless extra == REDUCE_NO_EXTRA) + the sum
// of partials [begin, begin + count), computed only when valid[gate] == 0:
// the sums that a per-item pairing check needs after its group check failed.
constexpr uint32_t REDUCE_NO_EXTRA = 0xffffffffu;
struct GatedReduce {
  uint32_t begin;
  uint32_t count;
  uint32_t extra;
  uint32_t gate;
  uint32_t out;
};

// e(points[a], PK) * e(points[b], -G2) == 1, and gate[item] must be set.
// e(points[a], Q_a) * e(+-points[b], Q_b) == 1 for item `item`: Q_a is the
// cached G2 point pk_slot (prepared lines), Q_b the cached G2 point pk_b or
// -G2 when pk_b == PAIR_NEG_G2; flags bit0 negates points[b].  flags bit1:
// the check belongs to an aggregated group whose combined check wrote
// valid[group]; when that is set the item is accepted without its own pairing.
constexpr uint32_t PAIR_NEG_G2 = 0xffffffffu;
constexpr uint32_t PAIR_NEG_B = 1, PAIR_GROUPED = 2;
struct PairJob {
  uint32_t item;
  uint32_t a;
  uint32_t b;
  uint32_t pk_slot;
  uint32_t pk_b;
  uint32_t flags;
  uint32_t group;
};

// Sign: e = H2S(sk || m_1..m_L || domain); A = B / (sk + e) computed as the
// comb MSM over scalars scaled by 1 / (sk + e).  Terms [term, term + L + 2).
struct SignJob {
  uint64_t sk_off;
  uint32_t item;
  uint32_t term;
  uint32_t L;
  uint32_t e_index;
  uint32_t a_point;   // reduce output holding A
  uint32_t pk_slot;
};

// ProofVerify challenge (proof.rs:2155-2297): compares H2S(c_arr) with c.
struct PvChallengeJob {
  uint64_t ph_off;
  uint32_t ph_len;
  uint32_t item;
  uint32_t R;
  uint32_t disc_begin;   // into disc_index[] / disc_scalar[]
  uint32_t abar, bbar, d, t1, t2;  // point indexes
  uint32_t dom_index;
  uint32_t c_index;
};

// ProofGen per item (proof.rs:1568-1645, 1754-1815, 2317-2377).
struct PgJob {
  uint64_t ph_off;
  uint64_t proof_off;     // output offset
  uint32_t ph_len;
  uint32_t item;
  uint32_t L, R, U;
  uint32_t disc_begin;    // disclosed indexes / scalar indexes
  uint32_t und_begin;     // undisclosed message scalar indexes
  uint32_t rs;            // first of 5 + U random scalars
  uint32_t e_index;
  uint32_t dom_index;
  uint32_t k_index;       // first of 7 derived scalars
  uint32_t a, b;          // points: A (signature), B
  uint32_t d, abar, bbar, t1, t2;  // reduce outputs
};

// ------------------------------------------------------------------ launchers
// (all asynchronous on the given stream)
void launch_fill_status(hipStream_t s, uint32_t n, uint8_t* status, uint8_t* gate, uint8_t gate_value);
// dst <- src, bytes; either end may be pinned host memory
void launch_copy(hipStream_t s, void* dst, const void* src, size_t bytes);
void launch_h2s(hipStream_t s, const H2sJob* jobs, uint32_t n, const uint8_t* blob,
                const uint8_t* dst, uint32_t dst_len, uint32_t* out_scalars);
void launch_hash_jobs(hipStream_t s, int kind, const H2sJob* jobs, uint32_t n, const uint8_t* blob,
                      const uint8_t* dst, uint32_t dst_len, uint8_t* out);
// message hashes (MsgHashJob; the b_1 / b_2 input words in `tail`, tail_blocks
// each; `zpad` the SHA-256 state after a zero block) and domains (DomainJob)
// in one launch
void launch_hash_stage(hipStream_t s, const DomainJob* dom, uint32_t n_dom, const MsgHashJob* msg,
                       uint32_t n_msg, const uint8_t* blob, const uint32_t* zpad, const uint32_t* tail,
                       uint32_t tail_blocks, const uint8_t* pk_bytes, const uint8_t* gen_comp,
                       uint32_t* out);
// the SHA-256 state after one all-zero block (8 words)
void launch_sha_zpad(hipStream_t s, uint32_t* out);
void launch_domain(hipStream_t s, const DomainJob* jobs, uint32_t n, const uint8_t* hdr_blob,
                   const uint8_t* pk_bytes, const uint8_t* gen_comp, uint32_t* out_scalars);
void launch_gen_seeds(hipStream_t s, uint32_t first, uint32_t count, uint32_t* v_state,
                      uint32_t* seeds_out);
void launch_gen_points(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* seeds,
                       uint32_t* gen_affine, uint8_t* gen_comp);
// scratch: count * COMB_WINDOWS * COMB_ENTRIES * COMB_ENTRY_WORDS words
void launch_comb_build(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* gen_affine,
                       uint32_t* tables, uint32_t* scratch);
// Wide tables of generators [first, first + count) from their 8-bit tables.
// scratch: count * COMB16_WINDOWS * COMB16_ENTRIES * COMB_ENTRY_WORDS words
void launch_comb16_build(hipStream_t s, uint32_t first, uint32_t count, const uint32_t* tables,
                         uint32_t* tables16, uint32_t* scratch);
// Terms whose generator slot is below n16 use the wide tables.  With
// gate_slot, chunk i is skipped when valid[gate_slot[i]] != 0.  With list,
// only the chunks list[0 .. *count) run (n: the list's capacity).
void launch_comb_msm(hipStream_t s, const CombChunk* chunks, uint32_t n, const uint32_t* term_gen,
                     const uint32_t* scalars, const uint32_t* tables, const uint32_t* tables16,
                     uint32_t n16, uint32_t* partials, const uint32_t* gate_slot = nullptr,
                     const uint8_t* valid = nullptr, const uint32_t* list = nullptr,
                     const uint32_t* count = nullptr);
// The live jobs of the gated launches that follow a level of group checks:
// job j of kind k (0 multi-term products, 1 comb chunks, 2 gated sums) is
// live when valid[gate[k][j]] == 0.  The select writes the live jobs' indices
// to list[k] (in no particular order) and their number to count[k]; count[]
// (3 words) must be zero before it runs.  The gated kernels then run over the
// lists, so their waves hold live jobs only.
struct GateLists {
  const uint32_t* gate[3];
  uint32_t n[3];
  uint32_t* list[3];
  uint32_t* count;
};
void launch_gate_select(hipStream_t s, const GateLists& g, const uint8_t* valid);
// groups[i] sums into points[out_point[i]]
void launch_reduce(hipStream_t s, const ReduceGroup* groups, const uint32_t* out_point, uint32_t n,
                   const uint32_t* partials, uint32_t* points, uint8_t* pinf);
// With list: only the jobs list[0 .. *count) run (selected live already).
void launch_reduce_gated(hipStream_t s, const GatedReduce* jobs, uint32_t n, const uint32_t* partials,
                         uint32_t* points, uint8_t* pinf, const uint8_t* valid,
                         const uint32_t* list = nullptr, const uint32_t* count = nullptr);
void launch_prepare_pk(hipStream_t s, const uint32_t* slots, uint32_t n, const uint8_t* pk_bytes,
                       uint8_t* pk_status, uint32_t* pk_lines);
void launch_prepare_neg_g2(hipStream_t s, uint32_t* lines);
// check_subgroup = false: launch_var_mul's extra lanes check the subgroup
void launch_decode_g1(hipStream_t s, const DecodeG1Job* jobs, uint32_t n, const uint8_t* blob,
                      uint32_t* points, uint8_t* pinf, uint8_t* status, bool check_subgroup = true);
void launch_decode_fr(hipStream_t s, const DecodeFrJob* jobs, uint32_t n, const uint8_t* blob,
                      uint32_t* scalars, uint8_t* status);
// ws: n * VAR_WS_WORDS words of device workspace.  n_sub > 0: also check
// the subgroup membership of the points of sub[].  sums: multi-term jobs
// (their workspace sum_ws), launched in the same grid before the products.
void launch_var_mul(hipStream_t s, const VarJob* jobs, uint32_t n, const uint32_t* points,
                    const uint8_t* pinf, const uint32_t* scalars, uint32_t* partials, uint32_t* ws,
                    const DecodeG1Job* sub = nullptr, uint32_t n_sub = 0, uint8_t* status = nullptr,
                    const VarSum* sums = nullptr, uint32_t n_sums = 0, uint32_t* sum_ws = nullptr,
                    const uint8_t* valid = nullptr, const uint32_t* sum_list = nullptr,
                    const uint32_t* sum_count = nullptr);
void launch_term_fill(hipStream_t s, const TermFill* jobs, uint32_t n, uint32_t* scalars,
                      const uint32_t* sum_pairs = nullptr);
// l2buf: PAIR_L2_WORDS words of device scratch per job, filled by
// launch_pair_lines and read by launch_pairing: per line and pairing the 9
// second factors of the sparse line product (68 x 2 x 9 x 14 words)
constexpr size_t PAIR_L2_WORDS = (size_t)68 * 2 * 9 * 14;
// partials[out[i]] = sum of the partials of groups[i] (Jacobian, no inversion)
void launch_reduce_partials(hipStream_t s, const ReduceGroup* groups, const uint32_t* out, uint32_t n,
                            uint32_t* partials);
void launch_compress(hipStream_t s, const CompressJob* jobs, uint32_t n, const uint32_t* points,
                     const uint8_t* pinf, uint8_t* out);
// The per-item checks that follow group checks: every job that needs no
// pairing (accepted through its group, or not live) gets its verdict here;
// the others are listed in list[0 .. *count) (count zeroed beforehand), so
// that the list-driven pair_lines / pairing launches below hold only checks
// that run.
void launch_pair_select(hipStream_t s, const PairJob* jobs, uint32_t n, const uint8_t* pk_status,
                        uint8_t* status, const uint8_t* gate, uint8_t* valid, uint32_t* list,
                        uint32_t* count);
// With list: job list[i] for i < *count (n bounds the grid), line blocks at
// slot i of l2buf.
void launch_pair_lines(hipStream_t s, const PairJob* jobs, uint32_t n, const uint32_t* points,
                       const uint8_t* pinf, const uint8_t* pk_status, const uint32_t* pk_lines,
                       const uint32_t* neg_g2_lines, uint32_t* l2buf,
                       const uint8_t* group_valid = nullptr, const uint32_t* list = nullptr,
                       const uint32_t* count = nullptr);
void launch_pairing(hipStream_t s, const PairJob* jobs, uint32_t n, const uint32_t* points,
                    const uint8_t* pinf, const uint8_t* pk_status, uint32_t* l2buf, uint8_t* status,
                    const uint8_t* gate, uint8_t* valid, const uint32_t* list = nullptr,
                    const uint32_t* count = nullptr);
void launch_sign_e(hipStream_t s, const SignJob* jobs, uint32_t n, const uint8_t* blob,
                   const uint8_t* pk_status, uint32_t* scalars, uint8_t* status);
void launch_sign_finalize(hipStream_t s, const SignJob* jobs, uint32_t n, const uint32_t* points,
                          const uint8_t* pinf, const uint32_t* scalars, uint8_t* status,
                          uint8_t* sig_out);
void launch_pv_challenge(hipStream_t s, const PvChallengeJob* jobs, uint32_t n,
                         const uint8_t* blob, const uint32_t* disc_index,
                         const uint32_t* disc_scalar, const uint32_t* points, const uint8_t* pinf,
                         const uint32_t* scalars, uint8_t* gate);
void launch_pg_chacha(hipStream_t s, const PgJob* jobs, uint32_t n, const uint8_t* seeds,
                      uint32_t* scalars);
void launch_pg_scalars(hipStream_t s, const PgJob* jobs, uint32_t n, uint32_t* scalars,
                       uint8_t* status);
void launch_pg_finalize(hipStream_t s, const PgJob* jobs, uint32_t n, const uint8_t* blob,
                        const uint32_t* disc_index, const uint32_t* disc_scalar,
                        const uint32_t* und_scalar, const uint32_t* points, const uint8_t* pinf,
                        uint32_t* scalars, uint8_t* status, uint8_t* proof_out);
void launch_g2_mul_gen(hipStream_t s, const uint8_t* sk, uint32_t n, uint8_t* pk_out, uint8_t* ok);

// ------------------------------------------------------------------ tuning
// Launch shapes picked by the measured autotuner (bbs_tune.cpp), per size
// class: Pippenger window width c and entries per piece S by terms per MSM,
// comb terms per lane by fixed-base terms per batch.  0 / false = no entry.
bool tuned_pip(uint32_t terms_per_msm, uint32_t& c, uint32_t& S);
uint32_t tuned_comb_chunk(uint64_t batch_terms);
// Whole-operation pipelines (kind 0 = Verify, 1 = ProofVerify), by items per
// batch: comb terms per lane, aggregation group size, groups per
// supergroup (0: none) and signatures per product job of a Verify group
// (0: two products per item), from the pipeline lines of the table (measured
// by running sets of such batches in flight).
bool tuned_pipeline(int kind, uint64_t items, uint32_t& comb_chunk, uint32_t& agg_group,
                    uint32_t& agg_super, uint32_t* agg_straus = nullptr);
// Shapes forced on the calling thread while the tuner times a candidate (or
// by cubbs_hip_test_force_shapes); 0 = not forced.
struct TuneForce {
  uint32_t pip_c = 0, pip_s = 0, comb_chunk = 0;
};
TuneForce& tune_force();

// Process-wide settings from the environment, read once at the first call
// (never per call: the entry points run on many threads at once, and a
// getenv racing a setenv is undefined).  0 = the library's own choice.
struct Knobs {
  uint32_t comb_chunk = 0;          // CUBBS_HIP_COMB_CHUNK: comb terms per lane
  uint32_t pippenger_min = 1024;    // CUBBS_HIP_PIPPENGER_MIN: item MSMs from this size
  uint32_t pip_window = 0;          // CUBBS_HIP_PIP_WINDOW (4..16)
  uint32_t pip_piece = 0;           // CUBBS_HIP_PIP_PIECE (1..256)
  uint32_t comb16_slots = 130;      // CUBBS_HIP_COMB16_GENERATORS
  uint32_t hw_queues = 4;           // GPU_MAX_HW_QUEUES (HIP's own setting)
  uint32_t pk_cache_entries = 4096; // CUBBS_HIP_PK_CACHE_ENTRIES: resident prepared keys
  uint64_t gen_table_budget = 32ull << 30;  // CUBBS_HIP_GEN_TABLE_BUDGET_BYTES: 8-bit comb tables
  uint32_t coalesce = 1;            // CUBBS_HIP_COALESCE: combine small concurrent primitive calls
  uint32_t coalesce_inflight = 2;   // CUBBS_HIP_COALESCE_INFLIGHT: combined calls at once
};
const Knobs& knobs();

// ------------------------------------------------------------------ Pippenger MSM
// One term of a large variable-base MSM: points[point] * scalars[scalar],
// summed into the output partial of MSM `job`.
struct PipTerm {
  uint32_t point;
  uint32_t scalar;
  uint32_t job;
};
// Shape shared by the Pippenger MSMs of a batch.  Each scalar is split into
// two 128-bit GLV halves recoded into W = ceil(129 / c) signed c-bit digits;
// a digit of magnitude m in window w goes to bucket m - 1 of window w, so a
// window has B = 2^(c-1) = 2^K buckets.  Buckets are cut into pieces of at
// most S entries.
struct PipShape {
  uint32_t c, W, B, K, S;
  uint32_t jobs;         // MSMs
  uint32_t terms;        // PipTerm count
  uint32_t nbt;          // jobs * W * B buckets
  uint32_t max_entries;  // 2 * terms * W (bucket entries)
  uint32_t max_pieces;   // max_entries / S + nbt + 1 (bound on sum ceil(count / S))
};
constexpr uint32_t PIP_MAX_TERMS = 1u << 26;
PipShape pip_shape(uint32_t terms, uint32_t jobs);
// window width minimising the FpM cost model for n terms per MSM
uint32_t pip_model_window(double terms_per_msm);
struct PipBuffers {
  uint32_t* cnt;      // [nbt + 1] entries per bucket
  uint32_t* off;      // [nbt + 1] first entry of each bucket
  uint32_t* cursor;   // [nbt] scatter cursors
  uint32_t* poff;     // [nbt + 1] first piece of each bucket
  uint32_t* pmap;     // [max_pieces] bucket of each piece
  uint32_t* sorted;   // [max_entries] term << 2 | half << 1 | negative
  uint32_t* hdr;      // [2] pieces, entries
  uint32_t* pieces;   // [max_pieces] Jacobian piece sums
  uint32_t* buckets;  // [nbt] Jacobian bucket sums, halved in place
};
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v);
void launch_pip_count(hipStream_t s, const PipShape& sh, const PipTerm* terms, const uint8_t* pinf,
                      const uint32_t* scalars, const PipBuffers& bf);
void launch_pip_scan(hipStream_t s, const PipShape& sh, const PipBuffers& bf);
void launch_pip_scatter(hipStream_t s, const PipShape& sh, const PipTerm* terms,
                        const uint8_t* pinf, const uint32_t* scalars, const PipBuffers& bf);
void launch_pip_accum(hipStream_t s, const PipShape& sh, const PipTerm* terms,
                      const uint32_t* points, const PipBuffers& bf);
void launch_pip_bucket(hipStream_t s, const PipShape& sh, const PipBuffers& bf);
void launch_pip_halve(hipStream_t s, const PipShape& sh, uint32_t step, const PipBuffers& bf);
void launch_pip_final(hipStream_t s, const PipShape& sh, const uint32_t* job_out,
                      const PipBuffers& bf, uint32_t* partials);

// test-only primitives
void launch_test_g1_mul(hipStream_t s, const uint8_t* pts, const uint8_t* scalars, uint32_t n,
                        uint8_t* out, uint8_t* ok);
void launch_test_pairing_gt(hipStream_t s, const uint8_t* g1, const uint8_t* g2, uint32_t n,
                            uint8_t* out, uint8_t* ok);
void launch_test_g2_decompress(hipStream_t s, const uint8_t* g2, uint32_t n, uint8_t* out,
                               uint8_t* ok);
void launch_fp_mul_bench(hipStream_t s, uint32_t* data, uint32_t n, uint32_t iters);
void launch_mad_peak(hipStream_t s, uint32_t* data, uint32_t n, uint32_t iters);

}  // namespace cubbs
