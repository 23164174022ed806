// Device-resident state shared by every entry point of the library:
// generator comb tables, prepared issuer public keys and -G2 lines.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
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
  uint32_t* gen_tables = nullptr;
  // wide (16-bit window) comb tables of slots [0, wide_slots); the buffer
  // holds wide_cap slots and is allocated once, never moved
  uint32_t* gen_tables16 = nullptr;
  uint32_t wide_cap = 0, wide_slots = 0;
  bool wide_init = false;
  uint8_t* gen_comp = nullptr;
  uint32_t* gen_vstate = nullptr;
  std::vector<uint8_t> gen_comp_host;                   // 48 bytes per slot
  std::unordered_map<std::string, uint32_t> gen_index;  // encoding -> slot
  // prepared public keys (status: 0 invalid, 1 valid, 2 identity)
  std::unordered_map<std::string, uint32_t> pk_slots;
  uint32_t pk_count = 0, pk_cap = 0;
  uint8_t* pk_bytes = nullptr;
  uint8_t* pk_status = nullptr;
  uint32_t* pk_lines = nullptr;
  uint32_t* neg_g2_lines = nullptr;
  // buffers replaced by a growth step stay allocated: batches that snapshot
  // them may still be in flight on other streams
  std::vector<void*> retired;
};

Context& ctx();
// The *_locked functions require ctx().mu to be held.
void init_locked(Context& c);
void ensure_generators_locked(Context& c, uint32_t slots);
void lookup_pks_locked(Context& c, const std::vector<const uint8_t*>& pks,
                       std::vector<uint32_t>& slots_out);
// Next work stream of the pool (round-robin); the pool owns it.
hipStream_t next_stream_locked(Context& c);

}  // namespace cubbs
