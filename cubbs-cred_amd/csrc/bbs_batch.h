// Batches of whole BBS operations staged on the device (see bbs_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_internal.h"
#include "host_ctx.h"

namespace cubbs {

// Batch kinds beyond the two public staged kinds of cubbs_hip_ext.h.
constexpr int KIND_SIGN = 2;
constexpr int KIND_PROOF_GEN = 3;
// A primitive program (the MSM / add / pairing entries): status slots instead
// of items, results assembled by the caller from h_status / h_valid / output.
constexpr int KIND_PRIM = 4;
// A Verify batch and a ProofVerify batch staged as one (every stage launches
// once for the items of both; results: the Verify items, then the others).
constexpr int KIND_VERIFY_SET = 5;

// One output point of a primitive program: its status slots (gst =
// PRIM_NO_SLOT when the row has no group slot), the offset of its 48-byte
// encoding in the output area and its result value (route id).
constexpr uint32_t PRIM_NO_SLOT = 0xffffffffu;
struct PrimRow {
  uint32_t st, gst;
  size_t off;
  uint8_t value;
};

struct KernelTimer {
  bool on = false;
  std::vector<std::string> names;
  std::vector<int> count;
  std::vector<hipEvent_t> starts, stops;
  std::vector<size_t> event_name;
};

// Reusable staging memory for repeated batches on one stream (a serving slot).
// Device and pinned host blocks are kept across batches; reset() recycles
// them once the previous batch on the stream has completed, so steady-state
// batches allocate nothing.
struct Arena {
  struct Block {
    uint8_t* p;
    size_t cap, used;
  };
  hipStream_t streams[4] = {};  // one per operation kind, so sub-batches overlap
  hipStream_t aux[4] = {};      // second stream of each kind (concurrent stages)
  hipEvent_t fork[4] = {}, join[4] = {};
  int cur = 0;                   // stream used by the next planned batch
  std::vector<Block> dev, host;
  size_t dev_bytes = 0, host_bytes = 0;  // held in blocks
  size_t min_block = 64ull << 20;
  uint8_t* alloc_dev(size_t bytes);
  uint8_t* alloc_host(size_t bytes);
  void reset();
  ~Arena();
};
// While set on a thread, batches planned on that thread take their memory and
// stream from the arena instead of allocating their own.
void set_thread_arena(Arena* a);
// Batches this thread prepares next run on one stream (no auxiliary stream
// for the decode / product branch): for callers that keep many batches in
// flight, where the fork's second queue costs more than it hides.
void set_thread_single_stream(bool on);
// Group size of the aggregated pairing checks of the Verify / ProofVerify
// batches this thread prepares next (0 = per-item checks only).
void set_thread_agg_group(uint32_t n);
void set_thread_agg_super(uint32_t n);
// Points per product job of an aggregated Verify group's A sides (0: the
// products c_i A_i and (-c_i e_i) A_i per item; 1..VAR_TERMS_MAX: dual
// multi-term jobs, and the items' own products gated after the group checks).
void set_thread_agg_straus(uint32_t n);

struct Batch {
  int kind = 0;
  hipStream_t stream = nullptr;
  bool from_arena = false;           // memory and streams belong to an Arena
  // Verify / ProofVerify run the point decode + variable-base multiplication
  // on `aux`, concurrently with hashing and the comb MSM on `stream`.
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  size_t n_total = 0;                // items passed in
  std::vector<int32_t> pre_code;     // per item: -1 = decided on the device
  std::vector<uint8_t> pre_value;    // value for items decided on the host
  std::vector<uint32_t> dev_index;   // per item: index among device items
  uint32_t nv = 0;                   // items on the device
  std::vector<size_t> out_off, out_len;  // ProofGen output slices per device item
  // staged inputs
  size_t in_bytes = 0, work_bytes = 0, out_bytes = 0;
  uint8_t* h_in = nullptr;
  uint8_t* d_in = nullptr;
  // pinned staging of the results: status, valid (each align256(slots)) and
  // the output area, written by a copy kernel and read after the stream sync
  uint8_t* h_res = nullptr;
  size_t res_stride = 0;
  uint8_t* d_work = nullptr;
  // job arrays (device)
  MsgHashJob* msg = nullptr;
  const uint32_t* tail = nullptr;     // b_1 / b_2 input words of the message hashes
  uint32_t tail_blocks = 0;
  const uint32_t* sha_zpad = nullptr;
  DomainJob* dom = nullptr;
  DecodeG1Job* dec_g1 = nullptr;
  DecodeFrJob* dec_fr = nullptr;
  TermFill* fills = nullptr;
  VarJob* var = nullptr;
  CombChunk* chunks = nullptr;
  uint32_t* term_gen = nullptr;
  ReduceGroup* groups1 = nullptr;
  uint32_t* out1 = nullptr;
  ReduceGroup* groups2 = nullptr;
  uint32_t* out2 = nullptr;
  PairJob* pairs = nullptr;
  // pairs[]: n_spairs supergroup checks, then n_gpairs group checks, then
  // the per-item checks
  uint32_t n_agg_slots = 0, n_gpairs = 0, n_spairs = 0;
  uint32_t agg_passed = 0;  // group checks that held at the last fetch of results
  std::vector<uint32_t> pair_group;   // checks after the supergroups': slot of the enclosing check (or PRIM_NO_SLOT)
  // summed term fills (TermFill mode 3) and the work gated on the group
  // checks: comb chunks and sums of the items of failing groups
  uint32_t* sum_pairs = nullptr;
  CombChunk* fb_chunks = nullptr;
  uint32_t* fb_gate = nullptr;
  GatedReduce* fb_red = nullptr;
  VarSum* fb_sum = nullptr;
  uint32_t n_fb_chunks = 0, n_fb_red = 0, n_fb_sum = 0;
  // gate slots of fb_sum and fb_red jobs (fb_gate: of fb_chunks), and the
  // lists of their live jobs that the select after the group checks writes
  // (products, chunks, sums; counts in gate_count[0..3))
  uint32_t* fb_sum_gate = nullptr;
  uint32_t* fb_red_gate = nullptr;
  uint32_t* gate_list = nullptr;
  uint32_t* gate_count = nullptr;
  // per gated job, for the work counts of a launch: its gate slot and the
  // comb terms, mixed additions and point additions it runs
  struct GatedWork {
    uint32_t gate, terms, madds, adds;
    uint32_t sums, sum_terms, duals, dual_terms;  // multi-term products (gated jobs)
  };
  std::vector<GatedWork> gated_work;
  uint64_t ungated_comb_terms = 0, ungated_comb_madds = 0, ungated_point_adds = 0;
  uint64_t ungated_straus[4] = {0, 0, 0, 0};  // sums, their terms, dual sums, their terms
  struct Level {
    ReduceGroup* groups;
    uint32_t* out;
    uint32_t n;
  };
  std::vector<Level> levels;          // partial-sum tree, in launch order
  // Pippenger MSMs (large variable-base MSMs of the primitive entries)
  PipShape pip{};
  PipTerm* pip_terms = nullptr;
  uint32_t* pip_out = nullptr;
  PipBuffers pipb{};
  CompressJob* comp = nullptr;
  uint32_t n_comp = 0;
  SignJob* sign = nullptr;
  PvChallengeJob* pvch = nullptr;
  PgJob* pg = nullptr;
  uint32_t* disc_index = nullptr;
  uint32_t* disc_scalar = nullptr;
  uint32_t* und_scalar = nullptr;
  uint8_t* map_dst = nullptr;
  uint32_t map_dst_len = 0;
  uint8_t* blob = nullptr;
  uint8_t* seeds = nullptr;
  bool seeded = false;
  uint32_t n_msg = 0, n_dom = 0, n_dec_g1 = 0, n_dec_fr = 0, n_fills = 0, n_var = 0, n_chunks = 0,
           n_terms = 0, n_groups1 = 0, n_groups2 = 0, n_pairs = 0, n_sign = 0, n_pvch = 0,
           n_pg = 0;
  uint64_t comb_terms = 0;
  uint64_t comb_madds = 0;            // mixed additions of the comb launch (at most)
  CubbsHipBatchWork work{};           // algorithmic work units (cubbs_hip_batch_work)
  // work arrays (device)
  uint32_t* scalars = nullptr;
  uint32_t* points = nullptr;
  uint8_t* pinf = nullptr;
  uint32_t* partials = nullptr;
  uint8_t* status = nullptr;
  uint8_t* gate = nullptr;
  uint8_t* valid = nullptr;
  uint8_t* out = nullptr;
  uint32_t* l2buf = nullptr;          // per pairing check: the line values -lambda x_P
  uint32_t* pair_list = nullptr;      // per-item checks that run after the group checks
  uint32_t* pair_count = nullptr;     // and their number
  uint32_t* vtab = nullptr;           // VAR_WS_WORDS per variable-base product job
  VarSum* var_sum = nullptr;          // multi-term product jobs
  uint32_t n_var_sum = 0;
  uint32_t* sum_ws = nullptr;         // their workspace
  uint32_t n_scalars = 0, n_points = 0, n_partials = 0;
  // snapshot of resident state
  const uint32_t* tables = nullptr;
  const uint32_t* tables16 = nullptr;
  uint32_t n16 = 0;                   // generator slots with wide comb tables
  const uint8_t* gen_comp = nullptr;
  const uint8_t* pk_bytes = nullptr;
  const uint8_t* pk_status = nullptr;
  const uint32_t* pk_lines = nullptr;
  const uint32_t* neg_g2_lines = nullptr;
  // host results
  std::vector<uint8_t> h_status, h_valid;
  // staged primitive programs: output rows and their collected encodings
  std::vector<PrimRow> rows;
  std::vector<uint8_t> h_out;
  KernelTimer timer;
  // the whole launch_batch() sequence captured once and replayed
  hipGraphExec_t graph = nullptr;
  // resident state the batch uses (pinned keys, epoch of its snapshot), and
  // private copies it owns when the resident tables cannot hold its needs
  ResidentHold hold;
  std::vector<void*> owned;
  ~Batch();
};

uint32_t comb_chunk_terms(uint64_t batch_terms = 0, bool tuned = false);
void fill_os_random(uint8_t* out, size_t n);

std::unique_ptr<Batch> plan_verify(const CuBbsBbsGpuVerifyItem* items, size_t n);
std::unique_ptr<Batch> plan_proof_verify(const CuBbsBbsGpuProofVerifyItem* items, size_t n);
std::unique_ptr<Batch> plan_verify_and_proof_verify(const CuBbsBbsGpuVerifyItem* vitems, size_t nv,
                                                    const CuBbsBbsGpuProofVerifyItem* pitems,
                                                    size_t np);
std::unique_ptr<Batch> plan_sign(const CuBbsBbsGpuSignItem* items, size_t n);
// mode 0: blinding from the OS CSPRNG, 1: per-item 32-byte seeds, 2: explicit scalars
std::unique_ptr<Batch> plan_proof_gen(const CuBbsBbsGpuProofGenItem* items, size_t n, int mode,
                                      const uint8_t* seeds32, const uint8_t* scalars_be,
                                      const size_t* scalar_offsets);
// Staged primitive programs (kinds CUBBS_HIP_BATCH_G1_MSM and
// CUBBS_HIP_BATCH_G1_MSM_COMMON_POINTS): validated and uploaded, not launched;
// nullptr with status set on a call-level error.
std::unique_ptr<Batch> stage_primitive(int kind, const void* items, size_t n, int32_t& status);
// One result per output row; encodings kept in b.h_out (48 bytes per row).
void collect_primitive(Batch& b, CuBbsBbsGpuResult* results);
void launch_batch(Batch& b);
// Copy status, valid and the output area into b.h_res and wait for the stream.
void fetch_results(Batch& b);
// Capture launch_batch() into b.graph (on private streams, so concurrent users
// of the pool streams are never recorded); b.stream then replays it.
void capture_batch(Batch& b);
void collect_batch(Batch& b, CuBbsBbsGpuResult* results, uint8_t* const* out_ptrs,
                   size_t* out_lens);

}  // namespace cubbs
