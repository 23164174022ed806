// Host side of libcubbs_bbs_hip.so: the C ABI v2 entry points, the
// device-resident state (generator comb tables, prepared issuer public keys,
// prepared -G2) and the batch planner that turns ABI items into the flat
// job arrays the kernels consume.
//
// Nothing here computes on the CPU: every curve, field and hash operation runs
// in the kernels of bbs_kernels.hip.  When no device is usable the entry points
// return CUBBS_BBS_GPU_ERR_INTERNAL (there is no CPU fallback).
#include <hip/hip_runtime.h>
#include <sys/random.h>

#include <cerrno>
#include <stdexcept>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_batch.h"
#include "host_ctx.h"

namespace cubbs {

static bool slice_ok(const CuBbsBbsGpuByteSlice& s) { return s.len == 0 || s.data != nullptr; }

struct Layout {
  size_t size = 0;
  size_t add(size_t bytes) {
    size_t off = size;
    size = align256(size + bytes);
    return off;
  }
};

void fill_os_random(uint8_t* out, size_t n) {
  size_t got = 0;
  while (got < n) {
    ssize_t r = getrandom(out + got, n - got, 0);
    if (r < 0 && errno == EINTR) continue;  // interrupted by a signal: try again
    if (r <= 0) throw std::runtime_error("getrandom failed");
    got += (size_t)r;
  }
}

// ------------------------------------------------------------------ settings
static uint64_t env_u64(const char* name, uint64_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  char* end = nullptr;
  const unsigned long long x = std::strtoull(v, &end, 10);
  return end == v ? dflt : (uint64_t)x;
}

const Knobs& knobs() {
  static const Knobs k = [] {
    Knobs d;
    d.comb_chunk = (uint32_t)std::min<uint64_t>(env_u64("CUBBS_HIP_COMB_CHUNK", 0), 64);
    d.pippenger_min = (uint32_t)std::min<uint64_t>(env_u64("CUBBS_HIP_PIPPENGER_MIN", 1024), 1u << 30);
    d.pip_window = (uint32_t)std::min<uint64_t>(env_u64("CUBBS_HIP_PIP_WINDOW", 0), 16);
    d.pip_piece = (uint32_t)std::min<uint64_t>(env_u64("CUBBS_HIP_PIP_PIECE", 0), 256);
    d.comb16_slots = (uint32_t)std::min<uint64_t>(env_u64("CUBBS_HIP_COMB16_GENERATORS", COMB16_DEFAULT_SLOTS), 4096);
    d.hw_queues = (uint32_t)std::min<uint64_t>(env_u64("GPU_MAX_HW_QUEUES", 4), 64);
    d.pk_cache_entries = (uint32_t)std::max<uint64_t>(16, std::min<uint64_t>(env_u64("CUBBS_HIP_PK_CACHE_ENTRIES", 4096), 1u << 20));
    d.gen_table_budget = env_u64("CUBBS_HIP_GEN_TABLE_BUDGET_BYTES", 32ull << 30);
    d.coalesce = env_u64("CUBBS_HIP_COALESCE", 1) != 0;
    d.coalesce_inflight = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(env_u64("CUBBS_HIP_COALESCE_INFLIGHT", 2), 16));
    return d;
  }();
  return k;
}

// ------------------------------------------------------------------ device context
Context& ctx() {
  static Context* c = new Context();
  return *c;
}

// Must be called with ctx().mu held.
void init_locked(Context& c) {
  if (c.ready) return;
  if (c.failed) throw GpuError{hipErrorNoDevice};
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    c.failed = true;
    throw GpuError{hipErrorNoDevice};
  }
  // Work streams: one per hardware queue of the process (HIP maps streams
  // onto GPU_MAX_HW_QUEUES queues, 4 unless configured), created together so
  // that they land on distinct queues; batches take them round-robin, so
  // concurrently running batches do not share a queue needlessly.
  {
    const int q = std::max(2, std::min((int)knobs().hw_queues, 32));
    c.pool.resize(q);
    for (auto& st : c.pool) check(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  }
  check(hipStreamCreateWithFlags(&c.setup, hipStreamNonBlocking));
  check(hipMalloc(&c.gen_vstate, 64));
  check(hipMalloc(&c.neg_g2_lines, PK_LINES_WORDS * 4));
  c.resident_bytes += 64 + PK_LINES_WORDS * 4;
  check(hipMalloc(&c.sha_zpad, 32));
  c.resident_bytes += 32;
  launch_sha_zpad(c.setup, c.sha_zpad);
  launch_prepare_neg_g2(c.setup, c.neg_g2_lines);
  check(hipGetLastError());
  check(hipStreamSynchronize(c.setup));
  c.ready = true;
}

hipStream_t next_stream_locked(Context& c) {
  hipStream_t s = c.pool[c.pool_next % c.pool.size()];
  c.pool_next++;
  return s;
}

// Wide comb tables for slots [0, min(slots, wide_cap)), after their 8-bit
// tables exist.  The capacity (CUBBS_HIP_COMB16_GENERATORS, default 130 =
// P1, Q1, H_1..H_128; 0 = none) is allocated at the first call and never
// reallocated, so batches that snapshot the pointer stay valid; when the
// allocation fails the library keeps to the 8-bit tables.  ctx().mu held.
static void ensure_wide_tables_locked(Context& c, uint32_t slots) {
  if (!c.wide_init) {
    c.wide_init = true;
    c.wide_cap = knobs().comb16_slots;
    if (c.wide_cap && hipMalloc(&c.gen_tables16, (size_t)c.wide_cap * COMB16_TABLE_WORDS * 4) != hipSuccess) {
      (void)hipGetLastError();
      c.gen_tables16 = nullptr;
      c.wide_cap = 0;
    }
    c.resident_bytes += (size_t)c.wide_cap * COMB16_TABLE_WORDS * 4;
  }
  // built from the slots' 8-bit tables
  const uint32_t upto = std::min({slots, c.wide_cap, c.table_slots});
  if (upto <= c.wide_slots) return;
  const uint32_t slice = 32;  // generators per build launch (1.9 GB of scratch)
  uint32_t* scratch = nullptr;
  const uint32_t k0 = std::min(slice, upto - c.wide_slots);
  if (hipMalloc(&scratch, (size_t)k0 * COMB16_WINDOWS * COMB16_ENTRIES * COMB_ENTRY_WORDS * 4) !=
      hipSuccess) {
    (void)hipGetLastError();  // no room to build now: these slots keep their 8-bit tables
    return;
  }
  for (uint32_t s = c.wide_slots; s < upto; s += slice) {
    const uint32_t k = std::min(slice, upto - s);
    launch_comb16_build(c.setup, s, k, c.gen_tables, c.gen_tables16, scratch);
    check(hipGetLastError());
  }
  check(hipStreamSynchronize(c.setup));
  check(hipFree(scratch));
  c.wide_slots = upto;
}

// ------------------------------------------------------------------ resident state
void retire_locked(Context& c, void* p, size_t bytes) {
  if (!p) return;
  c.retired.push_back(Context::Retired{p, c.epoch, bytes});
  c.epoch++;  // batches attached from now on cannot see p
}

static void free_retired_locked(Context& c) {
  const uint64_t oldest = c.live.empty() ? UINT64_MAX : c.live.begin()->first;
  size_t keep = 0;
  for (auto& r : c.retired) {
    if (r.epoch < oldest) {
      (void)hipFree(r.p);
      c.resident_bytes -= r.bytes;
    } else {
      c.retired[keep++] = r;
    }
  }
  c.retired.resize(keep);
}

void hold_locked(Context& c, ResidentHold& h) {
  if (h.held) return;
  h.held = true;
  h.epoch = c.epoch;
  c.live[h.epoch]++;
}

void release_locked(Context& c, ResidentHold& h) {
  if (!h.held) return;
  for (uint32_t s : h.pk_slots)
    if (s < c.pk_pins.size() && c.pk_pins[s]) c.pk_pins[s]--;
  h.pk_slots.clear();
  auto it = c.live.find(h.epoch);
  if (it != c.live.end() && --it->second == 0) c.live.erase(it);
  h.held = false;
  free_retired_locked(c);
}

template <class T>
static T* grow_copy(Context& c, T* old, size_t old_bytes, size_t new_bytes) {
  T* p = nullptr;
  check(hipMalloc(&p, new_bytes));
  c.resident_bytes += new_bytes;
  if (old && old_bytes) check(hipMemcpyAsync(p, old, old_bytes, hipMemcpyDeviceToDevice, c.setup));
  retire_locked(c, old, old_bytes);
  return p;
}

// Generator points and encodings of slots [0, slots); 8-bit comb tables of
// slots [0, min(slots, table_max)).  ctx().mu held.
void ensure_generators_locked(Context& c, uint32_t slots) {
  if (!c.table_max) {
    const uint64_t per = (uint64_t)COMB_TABLE_WORDS * 4;
    c.table_max = (uint32_t)std::max<uint64_t>(2, std::min<uint64_t>(knobs().gen_table_budget / per, 1u << 24));
  }
  if (slots > c.gen_slots) {
    if (slots > c.gen_cap) {
      const uint32_t cap = std::max<uint32_t>({slots, c.gen_cap * 2, 64u});
      c.gen_affine = grow_copy(c, c.gen_affine, (size_t)c.gen_slots * G1A_WORDS * 4, (size_t)cap * G1A_WORDS * 4);
      c.gen_comp = grow_copy(c, c.gen_comp, (size_t)c.gen_slots * 48, (size_t)cap * 48);
      c.gen_cap = cap;
    }
    const uint32_t first_chain = c.gen_slots == 0 ? 1 : c.gen_slots;
    const uint32_t count = slots - first_chain;
    if (count) {
      uint32_t* seeds = nullptr;
      check(hipMalloc(&seeds, (size_t)count * 12 * 4));
      launch_gen_seeds(c.setup, first_chain, count, c.gen_vstate, seeds);
      launch_gen_points(c.setup, first_chain, count, seeds, c.gen_affine, c.gen_comp);
      check(hipGetLastError());
      check(hipStreamSynchronize(c.setup));
      check(hipFree(seeds));
    }
    // host mirror of the encodings: recognises generator points in the
    // primitive MSM entries (their comb tables are resident)
    c.gen_comp_host.resize((size_t)slots * 48);
    check(hipMemcpy(c.gen_comp_host.data() + (size_t)c.gen_slots * 48, c.gen_comp + (size_t)c.gen_slots * 48,
                    (size_t)(slots - c.gen_slots) * 48, hipMemcpyDeviceToHost));
    for (uint32_t g = c.gen_slots; g < slots; g++)
      c.gen_index.emplace(std::string(reinterpret_cast<const char*>(c.gen_comp_host.data()) + 48 * g, 48), g);
    c.gen_slots = slots;
  }
  const uint32_t want = std::min(slots, c.table_max);
  if (want > c.table_slots) {
    const size_t per = (size_t)COMB_TABLE_WORDS * 4;
    if (want > c.table_cap) {
      const uint32_t cap = std::min<uint32_t>(c.table_max, std::max<uint32_t>({want, c.table_cap * 2, 64u}));
      c.gen_tables = grow_copy(c, c.gen_tables, (size_t)c.table_slots * per, (size_t)cap * per);
      c.table_cap = cap;
    }
    build_comb_tables(c, c.gen_tables, c.table_slots, want - c.table_slots);
    c.table_slots = want;
  }
  ensure_wide_tables_locked(c, slots);
}

// 8-bit comb tables of generator slots [first, first + count) into `tables`
// (indexed by slot), built on the setup stream in slices that bound the
// scratch buffer.  ctx().mu held.
void build_comb_tables(Context& c, uint32_t* tables, uint32_t first, uint32_t count) {
  if (!count) return;
  const uint32_t slice = 64;
  uint32_t* scratch = nullptr;
  check(hipMalloc(&scratch, (size_t)std::min(slice, count) * COMB_WINDOWS * COMB_ENTRIES * COMB_ENTRY_WORDS * 4));
  for (uint32_t s = 0; s < count; s += slice) {
    const uint32_t k = std::min(slice, count - s);
    launch_comb_build(c.setup, first + s, k, c.gen_affine, tables, scratch);
    check(hipGetLastError());
  }
  check(hipStreamSynchronize(c.setup));
  check(hipFree(scratch));
}

// Prepare public keys (decompression, subgroup check, Miller-loop lines) into
// the given slots of the arrays.  ctx().mu held.
void prepare_pks_locked(Context& c, const std::vector<const uint8_t*>& keys,
                        const std::vector<uint32_t>& slots, uint8_t* pk_bytes, uint8_t* pk_status,
                        uint32_t* pk_lines) {
  if (keys.empty()) return;
  std::vector<uint8_t> hb(keys.size() * 96);
  for (size_t i = 0; i < keys.size(); i++) std::memcpy(hb.data() + 96 * i, keys[i], 96);
  uint32_t* d_slots = nullptr;
  uint8_t* d_b = nullptr;
  check(hipMalloc(&d_slots, slots.size() * 4));
  check(hipMalloc(&d_b, hb.size()));
  check(hipMemcpyAsync(d_slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, c.setup));
  check(hipMemcpyAsync(d_b, hb.data(), hb.size(), hipMemcpyHostToDevice, c.setup));
  for (size_t i = 0; i < keys.size(); i++)  // scatter the bytes into their slots
    check(hipMemcpyAsync(pk_bytes + (size_t)slots[i] * 96, d_b + 96 * i, 96, hipMemcpyDeviceToDevice,
                         c.setup));
  launch_prepare_pk(c.setup, d_slots, (uint32_t)keys.size(), pk_bytes, pk_status, pk_lines);
  check(hipGetLastError());
  check(hipStreamSynchronize(c.setup));
  check(hipFree(d_slots));
  check(hipFree(d_b));
}

bool lookup_pks_locked(Context& c, const std::vector<const uint8_t*>& pks,
                       std::vector<uint32_t>& slots_out, ResidentHold& hold) {
  if (!c.pk_cap) {  // the table, allocated once
    const uint32_t cap = knobs().pk_cache_entries;
    check(hipMalloc(&c.pk_bytes, (size_t)cap * 96));
    check(hipMalloc(&c.pk_status, cap));
    check(hipMalloc(&c.pk_lines, (size_t)cap * PK_LINES_WORDS * 4));
    c.resident_bytes += (size_t)cap * (97 + PK_LINES_WORDS * 4);
    c.pk_key.assign(cap, std::string());
    c.pk_pins.assign(cap, 0);
    c.pk_used.assign(cap, 0);
    c.pk_cap = cap;
  }
  const uint64_t now = ++c.tick;
  slots_out.assign(pks.size(), 0);
  // distinct keys of the call: hits, and misses that need a slot
  std::unordered_map<std::string, uint32_t> seen;  // key -> index into slots_out's key list
  std::vector<std::string> miss;
  std::vector<int64_t> where(pks.size(), -1);       // hit slot, or -(1 + miss index)
  size_t distinct = 0;
  for (size_t i = 0; i < pks.size(); i++) {
    std::string key(reinterpret_cast<const char*>(pks[i]), 96);
    auto it = c.pk_slots.find(key);
    if (it != c.pk_slots.end()) {
      where[i] = it->second;
      if (seen.emplace(key, 0).second) distinct++;
      continue;
    }
    auto sm = seen.find(key);
    if (sm == seen.end()) {
      seen.emplace(key, (uint32_t)miss.size() + 1);
      distinct++;
      miss.push_back(key);
      where[i] = -(int64_t)miss.size();
    } else {
      where[i] = -(int64_t)sm->second;
    }
  }
  if (distinct > c.pk_cap) return false;
  // the hits are used now: a miss of this call must not take their slots
  for (size_t i = 0; i < pks.size(); i++)
    if (where[i] >= 0) c.pk_used[(uint32_t)where[i]] = now;
  // slots for the misses: free ones first, then the least recently used
  // entries that no live batch uses and this call does not hit
  std::vector<uint32_t> fresh;
  if (!miss.empty()) {
    std::vector<uint32_t> cand;
    for (uint32_t s = 0; s < c.pk_cap; s++)
      if (c.pk_key[s].empty() || (c.pk_pins[s] == 0 && c.pk_used[s] != now)) cand.push_back(s);
    if (cand.size() < miss.size()) return false;
    std::partial_sort(cand.begin(), cand.begin() + miss.size(), cand.end(), [&](uint32_t a, uint32_t b) {
      const bool fa = c.pk_key[a].empty(), fb = c.pk_key[b].empty();
      if (fa != fb) return fa;
      return c.pk_used[a] < c.pk_used[b];
    });
    fresh.assign(cand.begin(), cand.begin() + miss.size());
  }
  std::vector<const uint8_t*> fresh_bytes;
  for (size_t m = 0; m < miss.size(); m++) {
    const uint32_t s = fresh[m];
    if (!c.pk_key[s].empty()) c.pk_slots.erase(c.pk_key[s]);
    else c.pk_count++;
    c.pk_key[s] = miss[m];
    c.pk_slots[miss[m]] = s;
    c.pk_used[s] = now;
    fresh_bytes.push_back(reinterpret_cast<const uint8_t*>(c.pk_key[s].data()));
  }
  prepare_pks_locked(c, fresh_bytes, fresh, c.pk_bytes, c.pk_status, c.pk_lines);
  for (size_t i = 0; i < pks.size(); i++) {
    const uint32_t s = where[i] >= 0 ? (uint32_t)where[i] : fresh[(size_t)(-where[i] - 1)];
    slots_out[i] = s;
  }
  // pin each distinct slot once
  std::vector<uint32_t> uniq(slots_out);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  for (uint32_t s : uniq) {
    c.pk_pins[s]++;
    hold.pk_slots.push_back(s);
  }
  return true;
}

}  // namespace cubbs

using namespace cubbs;

// ==================================================================== C ABI
extern "C" {

int32_t cubbs_bbs_gpu_backend_info_v2(uint32_t abi_version, CuBbsBbsGpuBackendInfo* info_out) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || !info_out) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  info_out->backend_kind = CUBBS_BBS_GPU_BACKEND_KIND_CUDA;
  info_out->hardware_acceleration = 0;
  try {
    Context& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    init_locked(c);
    info_out->hardware_acceleration = 1;
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_bbs_gpu_verify_batch_v2(uint32_t abi_version, const CuBbsBbsGpuVerifyItem* items,
                                      size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!items || !results) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    auto b = plan_verify(items, item_count);
    launch_batch(*b);
    collect_batch(*b, results, nullptr, nullptr);
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_bbs_gpu_proof_verify_batch_v2(uint32_t abi_version,
                                            const CuBbsBbsGpuProofVerifyItem* items,
                                            size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!items || !results) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    auto b = plan_proof_verify(items, item_count);
    launch_batch(*b);
    collect_batch(*b, results, nullptr, nullptr);
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_bbs_gpu_sign_batch_v2(uint32_t abi_version, CuBbsBbsGpuSignItem* items,
                                    size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!items || !results) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    auto b = plan_sign(items, item_count);
    launch_batch(*b);
    std::vector<uint8_t*> outs(item_count);
    std::vector<size_t> lens(item_count, 0);
    for (size_t i = 0; i < item_count; i++) outs[i] = items[i].signature_out.data;
    collect_batch(*b, results, outs.data(), lens.data());
    for (size_t i = 0; i < item_count; i++)
      if (results[i].code == CUBBS_BBS_GPU_SUCCESS) items[i].signature_out.len = lens[i];
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

static int32_t proof_gen_common(uint32_t abi_version, CuBbsBbsGpuProofGenItem* items,
                                size_t item_count, int mode, const uint8_t* seeds,
                                const uint8_t* scalars, const size_t* offsets,
                                CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!items || !results) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    auto b = plan_proof_gen(items, item_count, mode, seeds, scalars, offsets);
    launch_batch(*b);
    std::vector<uint8_t*> outs(item_count);
    std::vector<size_t> lens(item_count, 0);
    for (size_t i = 0; i < item_count; i++) outs[i] = items[i].proof_out.data;
    collect_batch(*b, results, outs.data(), lens.data());
    for (size_t i = 0; i < item_count; i++)
      if (results[i].code == CUBBS_BBS_GPU_SUCCESS) items[i].proof_out.len = lens[i];
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_bbs_gpu_proof_gen_batch_v2(uint32_t abi_version, CuBbsBbsGpuProofGenItem* items,
                                         size_t item_count, CuBbsBbsGpuResult* results) {
  return proof_gen_common(abi_version, items, item_count, 0, nullptr, nullptr, nullptr, results);
}

int32_t cubbs_hip_proof_gen_seeded_batch(uint32_t abi_version, CuBbsBbsGpuProofGenItem* items,
                                         size_t item_count, const uint8_t* seeds32,
                                         CuBbsBbsGpuResult* results) {
  if (item_count && !seeds32) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return proof_gen_common(abi_version, items, item_count, 1, seeds32, nullptr, nullptr, results);
}

int32_t cubbs_hip_proof_gen_with_scalars_batch(uint32_t abi_version,
                                               CuBbsBbsGpuProofGenItem* items, size_t item_count,
                                               const uint8_t* scalars_be, const size_t* offsets,
                                               CuBbsBbsGpuResult* results) {
  if (item_count && (!scalars_be || !offsets)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return proof_gen_common(abi_version, items, item_count, 2, nullptr, scalars_be, offsets,
                          results);
}

// Applies prepare flags to the planning state of the calling thread for one
// prepare call (undone by the destructor).  kd: 0 Verify, 1 ProofVerify
// (the tuning table's size class of `items`), -1 other kinds.
namespace {
struct PrepareScope {
  bool ok = false;
  TuneForce saved;
  PrepareScope(int kd, bool whole_op, size_t items, uint32_t flags) : saved(tune_force()) {
    uint32_t group = (flags >> 8) & 15u;
    const uint32_t chunk = (flags >> 12) & 127u;
    uint32_t super = (flags >> 20) & 15u;
    uint32_t straus = (flags >> 24) & 15u;
    const bool agg = (flags & CUBBS_HIP_BATCH_AGGREGATE) != 0;
    if ((flags & ~(uint32_t)(CUBBS_HIP_BATCH_SINGLE_STREAM | CUBBS_HIP_BATCH_AGGREGATE |
                             CUBBS_HIP_BATCH_AGG_GROUP(15) | CUBBS_HIP_BATCH_COMB_CHUNK(127) |
                             CUBBS_HIP_BATCH_AGG_SUPER(15) | CUBBS_HIP_BATCH_AGG_STRAUS(15))) ||
        (group && (group < 2 || group > 8 || !agg)) || chunk > 64 || (chunk && !whole_op) ||
        (super && (super < 2 || super > 8 || !agg)) || (straus && (straus > VAR_TERMS_MAX || !agg)))
      return;
    if (agg) {
      // the tuned shape of the batch's size class where the flags leave it
      // open: groups of 8 and no supergroups without a table entry
      uint32_t c = 0, gsz = 0, sup = 0, st = 0;
      const bool tuned = tuned_pipeline(kd, items, c, gsz, sup, &st);
      if (!group) group = tuned && gsz ? gsz : 8u;
      if (!super && !(flags & CUBBS_HIP_BATCH_AGG_GROUP(15))) super = tuned ? sup : 0u;
      if (!straus && !(flags & CUBBS_HIP_BATCH_AGG_GROUP(15))) straus = tuned ? st : 0u;
    }
    if (chunk) tune_force().comb_chunk = chunk;
    set_thread_single_stream((flags & CUBBS_HIP_BATCH_SINGLE_STREAM) != 0);
    set_thread_agg_group(agg ? group : 0u);
    set_thread_agg_super(agg ? super : 0u);
    set_thread_agg_straus(straus);
    ok = true;
  }
  ~PrepareScope() {
    set_thread_single_stream(false);
    set_thread_agg_group(0);
    set_thread_agg_super(0);
    set_thread_agg_straus(0);
    tune_force() = saved;
  }
};
}  // namespace

void* cubbs_hip_batch_prepare_flags(uint32_t abi_version, int kind, const void* items,
                                    size_t item_count, uint32_t flags, int32_t* status_out) {
  int32_t dummy;
  int32_t& st = status_out ? *status_out : dummy;
  const int kd = kind == CUBBS_HIP_BATCH_VERIFY ? 0 : kind == CUBBS_HIP_BATCH_PROOF_VERIFY ? 1 : -1;
  PrepareScope scope(kd, kd >= 0, item_count, flags);
  if (!scope.ok) {
    st = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    return nullptr;
  }
  return cubbs_hip_batch_prepare(abi_version, kind, items, item_count, status_out);
}

void* cubbs_hip_batch_prepare_verify_set(uint32_t abi_version, const CuBbsBbsGpuVerifyItem* verify_items,
                                         size_t verify_count,
                                         const CuBbsBbsGpuProofVerifyItem* proof_items,
                                         size_t proof_count, uint32_t flags, int32_t* status_out) {
  int32_t dummy;
  int32_t& st = status_out ? *status_out : dummy;
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (verify_count && !verify_items) ||
      (proof_count && !proof_items)) {
    st = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    return nullptr;
  }
  PrepareScope scope(0, true, verify_count, flags);
  if (!scope.ok) {
    st = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    return nullptr;
  }
  try {
    std::unique_ptr<Batch> b = plan_verify_and_proof_verify(verify_items, verify_count, proof_items, proof_count);
    if (b->stream) check(hipStreamSynchronize(b->stream));
    st = CUBBS_BBS_GPU_SUCCESS;
    return b.release();
  } catch (...) {  // a GPU error, or host memory exhausted
    st = CUBBS_BBS_GPU_ERR_INTERNAL;
    return nullptr;
  }
}

void* cubbs_hip_batch_prepare(uint32_t abi_version, int kind, const void* items, size_t item_count,
                              int32_t* status_out) {
  int32_t dummy;
  int32_t& st = status_out ? *status_out : dummy;
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && !items)) {
    st = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    return nullptr;
  }
  try {
    std::unique_ptr<Batch> b;
    if (kind == CUBBS_HIP_BATCH_VERIFY) {
      b = plan_verify(static_cast<const CuBbsBbsGpuVerifyItem*>(items), item_count);
    } else if (kind == CUBBS_HIP_BATCH_PROOF_VERIFY) {
      b = plan_proof_verify(static_cast<const CuBbsBbsGpuProofVerifyItem*>(items), item_count);
    } else if (kind == CUBBS_HIP_BATCH_G1_MSM || kind == CUBBS_HIP_BATCH_G1_MSM_COMMON_POINTS) {
      b = stage_primitive(kind, items, item_count, st);
      if (!b) return nullptr;
    } else {
      st = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      return nullptr;
    }
    if (b->stream) check(hipStreamSynchronize(b->stream));
    st = CUBBS_BBS_GPU_SUCCESS;
    return b.release();
  } catch (...) {  // a GPU error, or host memory exhausted
    st = CUBBS_BBS_GPU_ERR_INTERNAL;
    return nullptr;
  }
}

int32_t cubbs_hip_batch_launch(void* batch) {
  if (!batch) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    Batch* b = static_cast<Batch*>(batch);
    if (b->graph)
      check(hipGraphLaunch(b->graph, b->stream));
    else
      launch_batch(*b);
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_hip_batch_sync(void* batch) {
  if (!batch) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  if (!b->stream) return CUBBS_BBS_GPU_SUCCESS;
  return hipStreamSynchronize(b->stream) == hipSuccess ? CUBBS_BBS_GPU_SUCCESS
                                                       : CUBBS_BBS_GPU_ERR_INTERNAL;
}

int32_t cubbs_hip_batch_results(void* batch, CuBbsBbsGpuResult* results) {
  if (!batch || !results) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    Batch* b = static_cast<Batch*>(batch);
    if (b->kind == KIND_PRIM)
      collect_primitive(*b, results);
    else
      collect_batch(*b, results, nullptr, nullptr);
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

void cubbs_hip_batch_free(void* batch) { delete static_cast<Batch*>(batch); }

int32_t cubbs_hip_batch_use_graph(void* batch, int32_t on) {
  if (!batch) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  try {
    if (on) {
      check(hipStreamSynchronize(b->stream));
      capture_batch(*b);
    } else if (b->graph) {
      check(hipStreamSynchronize(b->stream));
      check(hipGraphExecDestroy(b->graph));
      b->graph = nullptr;
    }
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_hip_batch_outputs(void* batch, uint8_t* out, size_t out_cap, size_t* out_len) {
  if (!batch || (out_cap && !out)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  if (b->kind != KIND_PRIM) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (out_cap < b->h_out.size()) return CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT;
  if (!b->h_out.empty()) std::memcpy(out, b->h_out.data(), b->h_out.size());
  if (out_len) *out_len = b->h_out.size();
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_batch_shape(void* batch, uint64_t* items, uint64_t* msm_terms,
                              uint64_t* comb_madds, uint64_t* pairings) {
  if (!batch) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  if (items) *items = b->nv;
  if (msm_terms) *msm_terms = b->comb_terms;
  if (comb_madds) *comb_madds = b->comb_madds;
  if (pairings) *pairings = 2ull * b->n_pairs;
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_resident_stats(CubbsHipResidentStats* out) {
  if (!out) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    Context& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    init_locked(c);
    CubbsHipResidentStats r{};
    r.resident_bytes = c.resident_bytes;
    r.pk_capacity = c.pk_cap;
    r.pk_entries = c.pk_count;
    for (uint32_t p : c.pk_pins) r.pk_pinned += p != 0;
    r.generator_slots = c.gen_slots;
    r.table_slots = c.table_slots;
    r.table_max_slots = c.table_max;
    r.wide_slots = c.wide_slots;
    r.retired_buffers = c.retired.size();
    for (const auto& e : c.live) r.live_batches += e.second;
    *out = r;
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_hip_batch_work(void* batch, CubbsHipBatchWork* out) {
  if (!batch || !out) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  *out = static_cast<Batch*>(batch)->work;
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_batch_group_checks(void* batch, uint64_t* groups, uint64_t* passed) {
  if (!batch) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  if (groups) *groups = b->n_agg_slots;
  if (passed) *passed = b->agg_passed;
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_batch_profile(void* batch, int iters, char* names_out, size_t names_cap,
                                double* ms_out, int max_kernels, double* total_ms_out) {
  if (!batch || iters <= 0) return -CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Batch* b = static_cast<Batch*>(batch);
  try {
    KernelTimer& t = b->timer;
    for (size_t i = 0; i < t.starts.size(); i++) {
      (void)hipEventDestroy(t.starts[i]);
      (void)hipEventDestroy(t.stops[i]);
    }
    t = KernelTimer();
    t.on = true;
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    check(hipEventRecord(e0, b->stream));
    for (int i = 0; i < iters; i++) launch_batch(*b);
    check(hipEventRecord(e1, b->stream));
    check(hipStreamSynchronize(b->stream));
    float total = 0;
    check(hipEventElapsedTime(&total, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (total_ms_out) *total_ms_out = total / iters;
    std::vector<double> sum(t.names.size(), 0.0);
    for (size_t k = 0; k < t.starts.size(); k++) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, t.starts[k], t.stops[k]));
      sum[t.event_name[k]] += ms;
    }
    int nk = (int)std::min<size_t>(t.names.size(), (size_t)max_kernels);
    std::string joined;
    for (int k = 0; k < nk; k++) {
      if (ms_out) ms_out[k] = sum[k] / std::max(1, t.count[k]);
      joined += t.names[k];
      joined += "\n";
    }
    if (names_out && names_cap) {
      size_t m = std::min(names_cap - 1, joined.size());
      std::memcpy(names_out, joined.data(), m);
      names_out[m] = 0;
    }
    t.on = false;
    return nk;
  } catch (...) {  // a GPU error, or host memory exhausted
    return -CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

// ---------------------------------------------------------------- hashing
// Shared implementation of the packed hash entries: kind 0 = SHA-256 digest,
// 1 = expand_message_xmd(.., 48), 2 = hash_to_scalar.
// Scalars of this thread's last hash_to_scalar / messages_to_scalar call: the
// "resident" message scalars of sign_e_from_resident_scalars.
static thread_local std::vector<uint8_t> tl_last_scalars;

static int32_t packed_hash(int kind, const uint8_t* packed, const size_t* offsets,
                           const size_t* lengths, size_t n, const uint8_t* dst, size_t dst_len,
                           uint8_t* out, bool record = true) {
  if (n == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!offsets || !lengths || !out) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (kind != 0 && (dst_len > 255 || (dst_len && !dst))) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  size_t total = 0;
  for (size_t i = 0; i < n; i++) {
    if (lengths[i] && !packed) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    total = std::max(total, offsets[i] + lengths[i]);
  }
  try {
    hipStream_t s;
    {
      Context& c = ctx();
      std::lock_guard<std::mutex> g(c.mu);
      init_locked(c);
      s = next_stream_locked(c);
    }
    const size_t out_w = kind == 0 ? 32 : (kind == 1 ? 48 : 32);
    Layout L;
    size_t o_jobs = L.add(sizeof(H2sJob) * n);
    size_t o_dst = L.add(std::max<size_t>(dst_len, 1));
    size_t o_blob = L.add(std::max<size_t>(total, 1));
    size_t o_out = L.add(out_w * n);
    std::vector<uint8_t> h(L.size);
    auto* jobs = reinterpret_cast<H2sJob*>(h.data() + o_jobs);
    for (size_t i = 0; i < n; i++) jobs[i] = H2sJob{offsets[i], (uint32_t)lengths[i], (uint32_t)i};
    if (dst_len) std::memcpy(h.data() + o_dst, dst, dst_len);
    if (total) std::memcpy(h.data() + o_blob, packed, total);
    uint8_t* d = nullptr;
    check(hipMalloc(&d, L.size));
    check(hipMemcpyAsync(d, h.data(), o_out, hipMemcpyHostToDevice, s));
    launch_hash_jobs(s, kind, reinterpret_cast<H2sJob*>(d + o_jobs), (uint32_t)n, d + o_blob,
                     d + o_dst, (uint32_t)dst_len, d + o_out);
    check(hipGetLastError());
    check(hipMemcpyAsync(out, d + o_out, out_w * n, hipMemcpyDeviceToHost, s));
    check(hipStreamSynchronize(s));
    check(hipFree(d));
    if (kind == 2 && record) tl_last_scalars.assign(out, out + 32 * n);
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_bbs_gpu_sha256_batch_for_test_v2(uint32_t abi_version, const uint8_t* packed,
                                               const size_t* offsets, const size_t* lengths,
                                               size_t item_count, uint8_t* digests_out) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return packed_hash(0, packed, offsets, lengths, item_count, nullptr, 0, digests_out);
}

int32_t cubbs_bbs_gpu_expand_xmd_sha256_48_batch_for_test_v2(
    uint32_t abi_version, const uint8_t* packed, const size_t* offsets, const size_t* lengths,
    size_t item_count, const uint8_t* dst, size_t dst_len, uint8_t* out48) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return packed_hash(1, packed, offsets, lengths, item_count, dst, dst_len, out48);
}

int32_t cubbs_bbs_gpu_hash_to_scalar_sha256_batch_for_test_v2(
    uint32_t abi_version, const uint8_t* packed, const size_t* offsets, const size_t* lengths,
    size_t item_count, const uint8_t* dst, size_t dst_len, uint8_t* scalars_out) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return packed_hash(2, packed, offsets, lengths, item_count, dst, dst_len, scalars_out);
}

// crates/cubbs-bbs/src/gpu_runtime.rs:763 ; reference semantics
// native/cuda/src/cubbs_bbs_cuda.cu:353-447 (one DST per call, 32-byte outputs).
int32_t cubbs_bbs_gpu_messages_to_scalar_batch_v2(uint32_t abi_version,
                                                  CuBbsBbsGpuMessagesToScalarItem* items,
                                                  size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  size_t total = 0;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (!slice_ok(it.message)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    if (it.dst.len > 255 || !slice_ok(it.dst) || !it.scalar_out.data || it.scalar_out.capacity < 32)
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    total += it.message.len;
  }
  const auto& d0 = items[0].dst;
  for (size_t i = 1; i < item_count; i++) {
    if (items[i].dst.len != d0.len ||
        (d0.len && std::memcmp(items[i].dst.data, d0.data, d0.len) != 0))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  std::vector<uint8_t> packed(std::max<size_t>(total, 1));
  std::vector<size_t> offs(item_count), lens(item_count);
  size_t cur = 0;
  for (size_t i = 0; i < item_count; i++) {
    offs[i] = cur;
    lens[i] = items[i].message.len;
    if (lens[i]) std::memcpy(packed.data() + cur, items[i].message.data, lens[i]);
    cur += lens[i];
  }
  std::vector<uint8_t> out(32 * item_count);
  int32_t st = packed_hash(2, packed.data(), offs.data(), lens.data(), item_count, d0.data, d0.len,
                           out.data());
  if (st != CUBBS_BBS_GPU_SUCCESS) return st;
  for (size_t i = 0; i < item_count; i++) {
    std::memcpy(items[i].scalar_out.data, out.data() + 32 * i, 32);
    items[i].scalar_out.len = 32;
    results[i].code = CUBBS_BBS_GPU_SUCCESS;
    results[i].value = 0;
    results[i].bytes_written = 32;
  }
  return CUBBS_BBS_GPU_SUCCESS;
}

// reference header :111-119, :254-257; semantics native/cuda/src/cubbs_bbs_cuda.cu:476-716
// (core_sign's e input, bbsplus/signature.rs:577: sk || m_1 .. m_L || domain).
int32_t cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuSignEFromResidentScalarsItem* items, size_t item_count,
    CuBbsBbsGpuByteSlice message_dst, CuBbsBbsGpuByteSlice e_dst, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || message_dst.len > 255 || e_dst.len > 255 ||
      !slice_ok(message_dst) || !slice_ok(e_dst) || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  const bool raw = items[0].messages != nullptr;
  size_t total = 0, raw_bytes = 0;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    const size_t count = it.message_count;
    if (count == 0 || it.message_scalar_offset + count < it.message_scalar_offset ||
        !it.secret_key.data || it.secret_key.len != 32 || !it.domain.data || it.domain.len != 32 ||
        !it.scalar_out.data || it.scalar_out.capacity < 32 || (it.messages != nullptr) != raw)
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    if (raw) {
      if (it.message_scalar_offset != total || !it.message_scalars_out.data ||
          it.message_scalars_out.capacity < count * 32)
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      for (size_t m = 0; m < count; m++) {
        if (!slice_ok(it.messages[m])) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
        raw_bytes += it.messages[m].len;
      }
      total += count;
    } else if (it.message_scalar_offset + count > tl_last_scalars.size() / 32) {
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    }
  }
  if (raw) {
    std::vector<uint8_t> packed(std::max<size_t>(raw_bytes, 1));
    std::vector<size_t> offs(total), lens(total);
    size_t k = 0, cur = 0;
    for (size_t i = 0; i < item_count; i++)
      for (size_t m = 0; m < items[i].message_count; m++, k++) {
        offs[k] = cur;
        lens[k] = items[i].messages[m].len;
        if (lens[k]) std::memcpy(packed.data() + cur, items[i].messages[m].data, lens[k]);
        cur += lens[k];
      }
    std::vector<uint8_t> sc(32 * total);
    const int32_t st = packed_hash(2, packed.data(), offs.data(), lens.data(), total,
                                   message_dst.data, message_dst.len, sc.data());
    if (st != CUBBS_BBS_GPU_SUCCESS) return st;
  }
  const std::vector<uint8_t>& scal = tl_last_scalars;
  std::vector<uint8_t> packed;
  std::vector<size_t> offs(item_count), lens(item_count);
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    offs[i] = packed.size();
    lens[i] = (it.message_count + 2) * 32;
    packed.insert(packed.end(), it.secret_key.data, it.secret_key.data + 32);
    const uint8_t* m = scal.data() + 32 * it.message_scalar_offset;
    packed.insert(packed.end(), m, m + 32 * it.message_count);
    packed.insert(packed.end(), it.domain.data, it.domain.data + 32);
  }
  std::vector<uint8_t> e(32 * item_count);
  const int32_t st = packed_hash(2, packed.data(), offs.data(), lens.data(), item_count, e_dst.data,
                                 e_dst.len, e.data(), false);
  if (st != CUBBS_BBS_GPU_SUCCESS) return st;
  for (size_t i = 0; i < item_count; i++) {
    auto& it = items[i];
    if (raw) {
      std::memcpy(it.message_scalars_out.data, scal.data() + 32 * it.message_scalar_offset,
                  32 * it.message_count);
      it.message_scalars_out.len = 32 * it.message_count;
    }
    std::memcpy(it.scalar_out.data, e.data() + 32 * i, 32);
    it.scalar_out.len = 32;
    results[i].code = CUBBS_BBS_GPU_SUCCESS;
    results[i].value = 0;
    results[i].bytes_written = 32;
  }
  return CUBBS_BBS_GPU_SUCCESS;
}

// ---------------------------------------------------------------- probes
}  // extern "C"
template <class F>
static int32_t run_probe(size_t in_bytes, const void* in, size_t out_bytes, void* out, size_t ok_bytes,
                         uint8_t* ok, F launch) {
  try {
    {
      Context& c = ctx();
      std::lock_guard<std::mutex> g(c.mu);
      init_locked(c);
    }
    hipStream_t s;
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint8_t* d = nullptr;
    size_t total = align256(in_bytes) + align256(out_bytes) + align256(ok_bytes) + 256;
    check(hipMalloc(&d, total));
    uint8_t* din = d;
    uint8_t* dout = d + align256(in_bytes);
    uint8_t* dok = dout + align256(out_bytes);
    check(hipMemsetAsync(d, 0, total, s));
    if (in_bytes) check(hipMemcpyAsync(din, in, in_bytes, hipMemcpyHostToDevice, s));
    launch(s, din, dout, dok);
    check(hipGetLastError());
    if (out_bytes) check(hipMemcpyAsync(out, dout, out_bytes, hipMemcpyDeviceToHost, s));
    if (ok_bytes) check(hipMemcpyAsync(ok, dok, ok_bytes, hipMemcpyDeviceToHost, s));
    check(hipStreamSynchronize(s));
    check(hipFree(d));
    check(hipStreamDestroy(s));
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

extern "C" {
int32_t cubbs_hip_test_g1_mul(const uint8_t* points48, const uint8_t* scalars32, size_t n,
                              uint8_t* out48, uint8_t* ok) {
  if (n == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!points48 || !scalars32 || !out48 || !ok) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  std::vector<uint8_t> in(80 * n);
  std::memcpy(in.data(), points48, 48 * n);
  std::memcpy(in.data() + 48 * n, scalars32, 32 * n);
  return run_probe(in.size(), in.data(), 48 * n, out48, n, ok,
                   [&](hipStream_t s, uint8_t* din, uint8_t* dout, uint8_t* dok) {
                     launch_test_g1_mul(s, din, din + 48 * n, (uint32_t)n, dout, dok);
                   });
}

int32_t cubbs_hip_test_pairing_gt(const uint8_t* g1_48, const uint8_t* g2_96, size_t n,
                                  uint8_t* out576, uint8_t* ok) {
  if (n == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!g1_48 || !g2_96 || !out576 || !ok) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  std::vector<uint8_t> in(144 * n);
  std::memcpy(in.data(), g1_48, 48 * n);
  std::memcpy(in.data() + 48 * n, g2_96, 96 * n);
  return run_probe(in.size(), in.data(), 576 * n, out576, n, ok,
                   [&](hipStream_t s, uint8_t* din, uint8_t* dout, uint8_t* dok) {
                     launch_test_pairing_gt(s, din, din + 48 * n, (uint32_t)n, dout, dok);
                   });
}

int32_t cubbs_hip_test_g2_roundtrip(const uint8_t* g2_96, size_t n, uint8_t* out96, uint8_t* ok) {
  if (n == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!g2_96 || !out96 || !ok) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return run_probe(96 * n, g2_96, 96 * n, out96, n, ok,
                   [&](hipStream_t s, uint8_t* din, uint8_t* dout, uint8_t* dok) {
                     launch_test_g2_decompress(s, din, (uint32_t)n, dout, dok);
                   });
}

int32_t cubbs_hip_generators(size_t count, uint8_t* out48) {
  if (count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!out48) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    Context& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    init_locked(c);
    ensure_generators_locked(c, (uint32_t)count);
    check(hipMemcpy(out48, c.gen_comp, 48 * count, hipMemcpyDeviceToHost));
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

static int32_t time_probe(uint32_t lanes, uint32_t iters, double* ms,
                          void (*launch)(hipStream_t, uint32_t*, uint32_t, uint32_t)) {
  if (!lanes || !ms) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    {
      Context& c = ctx();
      std::lock_guard<std::mutex> g(c.mu);
      init_locked(c);
    }
    hipStream_t s;
    check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    uint32_t* d = nullptr;
    // two Fp operands per lane (2 x 14 limbs of 28 bits), values below 2^380 < p
    check(hipMalloc(&d, (size_t)lanes * 28 * 4));
    std::vector<uint32_t> h((size_t)lanes * 28);
    for (size_t i = 0; i < h.size(); i++)
      h[i] = (uint32_t)(i * 2654435761u) & (i % 14 == 13 ? 0xffffu : 0x0fffffffu);
    check(hipMemcpyAsync(d, h.data(), h.size() * 4, hipMemcpyHostToDevice, s));
    launch(s, d, lanes, 1);  // warm-up
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    check(hipEventRecord(e0, s));
    launch(s, d, lanes, iters);
    check(hipEventRecord(e1, s));
    check(hipStreamSynchronize(s));
    float t = 0;
    check(hipEventElapsedTime(&t, e0, e1));
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check(hipFree(d));
    check(hipStreamDestroy(s));
    return CUBBS_BBS_GPU_SUCCESS;
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_hip_fp_mul_bench(uint32_t lanes, uint32_t iters, double* ms) {
  return time_probe(lanes, iters, ms, launch_fp_mul_bench);
}

int32_t cubbs_hip_mad_peak_bench(uint32_t lanes, uint32_t iters, double* ms) {
  return time_probe(lanes, iters, ms, launch_mad_peak);
}

int32_t cubbs_hip_sk_to_pk(const uint8_t* sk32, size_t n, uint8_t* pk96, uint8_t* ok) {
  if (n == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (!sk32 || !pk96 || !ok) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return run_probe(32 * n, sk32, 96 * n, pk96, n, ok,
                   [&](hipStream_t s, uint8_t* din, uint8_t* dout, uint8_t* dok) {
                     launch_g2_mul_gen(s, din, (uint32_t)n, dout, dok);
                   });
}

}  // extern "C"
