// Device-resident state shared by every entry point of the library:
// generator comb tables, prepared issuer public keys and -G2 lines.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "bbs_internal.h"

namespace cubbs {

struct GpuError {
  hipError_t err;
};
inline void check(hipError_t e) {
  if (e != hipSuccess) throw GpuError{e};
}
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Resident state is bounded: the prepared public keys live in a fixed table
// of pk_cap entries (CUBBS_HIP_PK_CACHE_ENTRIES, allocated once), evicted
// least recently used among the entries no live batch references; the 8-bit
// generator comb tables grow up to CUBBS_HIP_GEN_TABLE_BUDGET_BYTES.  A batch
// that needs more keys or generator tables than that holds private copies,
// freed with the batch.  Buffers replaced by a growth step are freed once no
// batch attached before the replacement is alive.
struct Context {
  std::mutex mu;
  bool ready = false;
  bool failed = false;
  hipStream_t setup = nullptr;
  std::vector<hipStream_t> pool;  // work streams, one per hardware queue
  uint32_t pool_next = 0;
  // generators: slot 0 = P1, slot 1 = Q1, slot 1 + j = H_j
  uint32_t gen_slots = 0, gen_cap = 0;
  uint32_t* gen_affine = nullptr;
  uint8_t* gen_comp = nullptr;
  uint32_t* gen_vstate = nullptr;
  uint32_t* sha_zpad = nullptr;  // SHA-256 state after one zero block (8 words)
  // 8-bit comb tables of slots [0, table_slots), capacity table_cap slots,
  // never more than table_max (the budget)
  uint32_t* gen_tables = nullptr;
  uint32_t table_slots = 0, table_cap = 0, table_max = 0;
  // wide (16-bit window) comb tables of slots [0, wide_slots); the buffer
  // holds wide_cap slots and is allocated once, never moved
  uint32_t* gen_tables16 = nullptr;
  uint32_t wide_cap = 0, wide_slots = 0;
  bool wide_init = false;
  std::vector<uint8_t> gen_comp_host;                   // 48 bytes per slot
  std::unordered_map<std::string, uint32_t> gen_index;  // encoding -> slot
  // prepared public keys (status: 0 invalid, 1 valid, 2 identity)
  std::unordered_map<std::string, uint32_t> pk_slots;  // key bytes -> slot
  std::vector<std::string> pk_key;                     // slot -> key bytes ("" = free)
  std::vector<uint32_t> pk_pins;                       // live batches using the slot
  std::vector<uint64_t> pk_used;                       // last lookup (tick)
  uint64_t tick = 0;
  uint32_t pk_cap = 0, pk_count = 0;
  uint8_t* pk_bytes = nullptr;
  uint8_t* pk_status = nullptr;
  uint32_t* pk_lines = nullptr;
  uint32_t* neg_g2_lines = nullptr;
  // buffers replaced by a growth step at epoch e: referenced at most by
  // batches attached at epochs <= e; freed when none of those is alive
  struct Retired {
    void* p;
    uint64_t epoch;
    size_t bytes;
  };
  std::vector<Retired> retired;
  uint64_t epoch = 0;
  std::map<uint64_t, uint32_t> live;  // epoch -> live batches attached then
  size_t resident_bytes = 0;          // device bytes of all of the above
};

// A batch's hold on resident state (taken by attach, released by ~Batch).
struct ResidentHold {
  bool held = false;
  uint64_t epoch = 0;
  std::vector<uint32_t> pk_slots;  // pinned cache slots
};

Context& ctx();
// The *_locked functions require ctx().mu to be held.
void init_locked(Context& c);
// Generator points (and encodings) of slots [0, slots); their 8-bit tables as
// far as the budget allows (c.table_slots).
void ensure_generators_locked(Context& c, uint32_t slots);
// Map public keys to cache slots, preparing unseen ones, and pin the slots in
// `hold`.  Returns false (nothing pinned) when the table cannot hold all the
// distinct keys at once; the caller then prepares private copies.
bool lookup_pks_locked(Context& c, const std::vector<const uint8_t*>& pks,
                       std::vector<uint32_t>& slots_out, ResidentHold& hold);
// 8-bit comb tables of generator slots [first, first + count) into `tables`
// (indexed by slot); synchronous on the setup stream.
void build_comb_tables(Context& c, uint32_t* tables, uint32_t first, uint32_t count);
// Prepared public keys (decompression, subgroup check, lines) of `keys` into
// `slots` of the given arrays; synchronous on the setup stream.
void prepare_pks_locked(Context& c, const std::vector<const uint8_t*>& keys,
                        const std::vector<uint32_t>& slots, uint8_t* pk_bytes, uint8_t* pk_status,
                        uint32_t* pk_lines);
// Start / end a batch's hold (epoch for retired buffers, pinned keys).
void hold_locked(Context& c, ResidentHold& hold);
void release_locked(Context& c, ResidentHold& hold);
// Replace a grown buffer: the old one is freed once no batch can use it.
void retire_locked(Context& c, void* p, size_t bytes);
// Next work stream of the pool (round-robin); the pool owns it.
hipStream_t next_stream_locked(Context& c);

}  // namespace cubbs
