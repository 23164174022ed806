// The reference ABI's primitive entry points (native/cuda/include/
// cubbs_bbs_gpu_abi.h:111-225, 254-308): G1 multi-scalar multiplications in
// several shapes, G1 addition, pairing equalities, prepare_b, sign e and the
// ProofGen T2 preparation.
//
// Every entry compiles its items into one device program (a KIND_PRIM batch
// of bbs_batch.cpp): compressed points are decoded with the subgroup check,
// 32-byte big-endian scalars decoded (canonical check), each product
// k * P runs as a one-lane windowed GLV multiplication, or as comb terms over the
// resident tables when the points are the ciphersuite generators, partial
// sums are added in a parallel tree (large MSMs) and the results compressed.
// Pairing equalities use the prepared-line cache and the 12-lane pairing.
//
// Error convention (uniform over the entries): structural errors (lengths,
// NULL pointers, output capacity) fail the call with
// CUBBS_BBS_GPU_ERR_INVALID_INPUT; an undecodable point or non-canonical
// scalar marks the affected row with code 1; a device failure returns
// CUBBS_BBS_GPU_ERR_INTERNAL.  result.value carries the reference's route id
// for MSM shapes (1 item, 2 common points, 3 large, 6 dual) and the boolean
// of a pairing equality.
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_batch.h"
#include "bbs_plan.h"
#include "host_ctx.h"

namespace cubbs {
namespace {

constexpr size_t LARGE_TERMS = 1024; // route id 3 from this many terms on

// Item MSMs of at least this many terms run as Pippenger bucket MSMs
// (CUBBS_HIP_PIPPENGER_MIN, default LARGE_TERMS), smaller ones as one-lane
// windowed GLV products summed in the partial-sum tree.
size_t pippenger_min() { return knobs().pippenger_min ? knobs().pippenger_min : LARGE_TERMS; }
constexpr uint8_t ROUTE_ITEM = CUBBS_BBS_MSM_ROUTE_CUBBS_ITEM;
constexpr uint8_t ROUTE_COMMON = CUBBS_BBS_MSM_ROUTE_CUBBS_COMMON_POINTS;
constexpr uint8_t ROUTE_LARGE = CUBBS_BBS_MSM_ROUTE_CUBBS_LARGE;
constexpr uint8_t ROUTE_DUAL = 6;  // the value the reference reports for dual MSMs

bool fixed(const CuBbsBbsGpuByteSlice& s, size_t n) { return s.len == n && s.data != nullptr; }
bool outcap(const CuBbsBbsGpuByteBuffer& b, size_t n) { return b.data != nullptr && b.capacity >= n; }

// A device program under construction.
struct Prog {
  Plan p;
  uint32_t slots = 0;                       // status slots
  std::vector<const uint8_t*> g2;           // G2 points to make resident (pairings)
  uint32_t max_gen = 2;                     // generators needed by comb terms
  uint64_t comb_terms = 0;                  // fixed-base terms planned (tuning class)

  uint32_t slot() { return slots++; }
  uint32_t point(const uint8_t* c48, uint32_t st) {
    const uint32_t pt = p.points(1);
    p.dec_g1.push_back(DecodeG1Job{p.bytes(c48, 48), pt, st});
    return pt;
  }
  // flags as DecodeFrJob (bit0 negate, bit1 reduce mod r)
  void scalar_at(uint32_t sc, const uint8_t* be32, uint32_t st, uint32_t flags) {
    p.dec_fr.push_back(DecodeFrJob{p.bytes(be32, 32), sc, st, flags, 0});
  }
  uint32_t scalar(const uint8_t* be32, uint32_t st, uint32_t flags = 0) {
    const uint32_t sc = p.scalars(1);
    scalar_at(sc, be32, st, flags);
    return sc;
  }
  void product(uint32_t pt, uint32_t sc) {
    p.var.push_back(VarJob{pt, sc, p.partials(VAR_PARTIALS)});
  }
  // Comb terms: scalars [sc, sc + n) against generator slots gen[0..n).
  void comb(uint32_t sc, const std::vector<uint32_t>& gen) {
    if (p.term_gen.size() < sc + gen.size()) p.term_gen.resize(sc + gen.size());
    for (size_t t = 0; t < gen.size(); t++) {
      p.term_gen[sc + t] = gen[t];
      max_gen = std::max(max_gen, gen[t] + 1);
    }
    p.comb(sc, (uint32_t)gen.size(), comb_chunk_terms(comb_terms, true));
  }
  // Sum of the partials [first, n_partials) into a new affine point slot.
  uint32_t finish(uint32_t first) {
    const uint32_t pt = p.points(1);
    p.reduce_to(first, p.n_partials - first, pt);
    return pt;
  }
  size_t compress(uint32_t pt) {
    const size_t off = p.out_bytes;
    p.out_bytes += 48;
    p.comp.push_back(CompressJob{pt, 0, off});
    return off;
  }
};

// Plans a program onto the device (inputs uploaded, nothing launched).
std::unique_ptr<Batch> stage(Prog& g) {
  auto b = std::make_unique<Batch>();
  b->kind = KIND_PRIM;
  b->nv = std::max<uint32_t>(g.slots, 1);
  std::vector<uint32_t> pk_slot;
  attach(*b, g.max_gen, g.g2, pk_slot);
  for (auto& pj : g.p.pairs) {  // pk_slot / pk_b hold indexes into g.g2 until here
    pj.pk_slot = pk_slot[pj.pk_slot];
    if (pj.pk_b != PAIR_NEG_G2) pj.pk_b = pk_slot[pj.pk_b];
  }
  for (auto& d : g.p.dom) d.pk_slot = pk_slot[d.pk_slot];  // the same for domains
  if (g.p.term_gen.size() < g.p.n_scalars) g.p.term_gen.resize(g.p.n_scalars);
  materialize(*b, g.p);
  return b;
}

// Status / valid / output bytes of a launched program, on the host.
struct Ran {
  std::vector<uint8_t> status, valid, out;
};
Ran fetch(Batch& b) {
  fetch_results(b);
  Ran r;
  r.status.assign(b.h_res, b.h_res + b.nv);
  r.valid.assign(b.h_res + b.res_stride, b.h_res + b.res_stride + b.nv);
  r.out.assign(b.h_res + 2 * b.res_stride, b.h_res + 2 * b.res_stride + std::max<size_t>(b.out_bytes, 1));
  return r;
}

// Runs a program: plan -> device -> status / valid / output bytes on the host.
Ran run(Prog& g) {
  auto b = stage(g);
  launch_batch(*b);
  return fetch(*b);
}

void set(CuBbsBbsGpuResult& r, int32_t code, uint8_t value, size_t bytes) {
  r.code = code;
  r.value = value;
  r.bytes_written = bytes;
}

// Generator slots of a point list when every point is a ciphersuite
// generator (then the resident comb tables serve the products).
bool as_generators(const CuBbsBbsGpuByteSlice* pts, size_t n, std::vector<uint32_t>& gen) {
  if (n == 0) return false;
  Context& c = ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  init_locked(c);
  // P1 first and a contiguous prefix is the common BBS shape: make sure the
  // generators up to that length exist before matching
  if (c.gen_slots < n && n < 4096 &&
      c.gen_index.count(std::string(reinterpret_cast<const char*>(pts[0].data), 48)))
    ensure_generators_locked(c, (uint32_t)n);
  gen.resize(n);
  for (size_t i = 0; i < n; i++) {
    auto it = c.gen_index.find(std::string(reinterpret_cast<const char*>(pts[i].data), 48));
    if (it == c.gen_index.end()) return false;
    gen[i] = it->second;
  }
  return true;
}

// One output row of a common-points shape: sum_t scalar(t) * P_t.
// `pts` are the decoded common points (var path) or empty (comb path with gen).
template <class ScalarAt>
void common_row(Prog& g, const std::vector<uint32_t>& pts, const std::vector<uint32_t>& gen,
                size_t terms, uint32_t st, ScalarAt scalar_at) {
  if (!gen.empty()) {
    const uint32_t sc = g.p.scalars((uint32_t)terms);
    for (size_t t = 0; t < terms; t++) g.scalar_at(sc + (uint32_t)t, scalar_at(t), st, 0);
    g.comb(sc, gen);
  } else {
    for (size_t t = 0; t < terms; t++) g.product(pts[t], g.scalar(scalar_at(t), st));
  }
}

// Result of one output point: code 1 when a status slot of the row is set,
// else success with the route id and the 48-byte point copied to `buf`.
void row_result(const PrimRow& row, const uint8_t* status, const uint8_t* out,
                CuBbsBbsGpuResult& res, CuBbsBbsGpuByteBuffer* buf) {
  if (status[row.st] || (row.gst != PRIM_NO_SLOT && status[row.gst])) {
    set(res, CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
    return;
  }
  if (buf) {
    std::memcpy(buf->data, out + row.off, 48);
    buf->len = 48;
  }
  set(res, CUBBS_BBS_GPU_SUCCESS, row.value, 48);
}

// Structure of item MSMs: 48-byte points and 32-byte scalars (call level).
bool msm_items_ok(const CuBbsBbsGpuG1MsmItem* items, size_t n) {
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    if (it.term_count && (!it.points || !it.scalars)) return false;
    for (size_t t = 0; t < it.term_count; t++)
      if (!fixed(it.points[t], 48) || !fixed(it.scalars[t], 32)) return false;
  }
  return true;
}

// One row per item: Pippenger from pippenger_min() terms on, else GLV
// products summed in the partial-sum tree.
void compile_msm(Prog& g, const CuBbsBbsGpuG1MsmItem* items, size_t n, std::vector<PrimRow>& rows) {
  const size_t pip_min = pippenger_min();
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    const uint32_t st = g.slot();
    const uint32_t first = g.p.n_partials;
    if (it.term_count >= pip_min && g.p.pip_terms.size() + it.term_count <= PIP_MAX_TERMS) {
      const uint32_t job = g.p.pip_job();
      for (size_t t = 0; t < it.term_count; t++)
        g.p.pip_terms.push_back(PipTerm{g.point(it.points[t].data, st),
                                        g.scalar(it.scalars[t].data, st), job});
    } else {
      for (size_t t = 0; t < it.term_count; t++)
        g.product(g.point(it.points[t].data, st), g.scalar(it.scalars[t].data, st));
    }
    rows.push_back(PrimRow{st, PRIM_NO_SLOT, g.compress(g.finish(first)),
                           it.term_count >= LARGE_TERMS ? ROUTE_LARGE : ROUTE_ITEM});
  }
}

// Structure of common-point groups (call level; output buffers not checked).
bool common_items_ok(const CuBbsBbsGpuG1MsmCommonPointsBatchItem* items, size_t n) {
  for (size_t gi = 0; gi < n; gi++) {
    const auto& it = items[gi];
    const bool packed = it.scalar_rows_be != nullptr;
    if (it.term_count > (SIZE_MAX / 32u) - 1u ||
        (it.term_count && it.output_count > SIZE_MAX / it.term_count) ||
        (it.term_count && !it.points) ||
        (it.output_count && !packed && !it.scalar_rows) ||
        (packed && it.scalar_row_stride < it.term_count * 32u))
      return false;
    for (size_t t = 0; t < it.term_count; t++)
      if (!fixed(it.points[t], 48)) return false;
    if (!packed)
      for (size_t r = 0; r < it.output_count; r++)
        for (size_t t = 0; t < it.term_count; t++)
          if (!fixed(it.scalar_rows[r * it.term_count + t], 32)) return false;
  }
  return true;
}

// One row per output of each group, in group order.
void compile_common(Prog& g, const CuBbsBbsGpuG1MsmCommonPointsBatchItem* items, size_t n,
                    std::vector<PrimRow>& rows) {
  for (size_t gi = 0; gi < n; gi++) g.comb_terms += (uint64_t)items[gi].output_count * items[gi].term_count;
  for (size_t gi = 0; gi < n; gi++) {
    const auto& it = items[gi];
    const uint32_t gst = g.slot();
    std::vector<uint32_t> gen, pts;
    if (!as_generators(it.points, it.term_count, gen)) {
      gen.clear();
      for (size_t t = 0; t < it.term_count; t++) pts.push_back(g.point(it.points[t].data, gst));
    }
    for (size_t r = 0; r < it.output_count; r++) {
      const uint32_t st = g.slot();
      const uint32_t first = g.p.n_partials;
      auto sc = [&](size_t t) {
        return it.scalar_rows_be ? it.scalar_rows_be + r * it.scalar_row_stride + 32 * t
                                 : it.scalar_rows[r * it.term_count + t].data;
      };
      common_row(g, pts, gen, it.term_count, st, sc);
      rows.push_back(PrimRow{st, gst, g.compress(g.finish(first)), ROUTE_COMMON});
    }
  }
}

int32_t guard(const std::function<int32_t()>& f) {
  try {
    return f();
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

}  // namespace

std::unique_ptr<Batch> stage_primitive(int kind, const void* items, size_t n, int32_t& status) {
  Prog g;
  std::vector<PrimRow> rows;
  if (kind == CUBBS_HIP_BATCH_G1_MSM) {
    auto* it = static_cast<const CuBbsBbsGpuG1MsmItem*>(items);
    if (!msm_items_ok(it, n)) {
      status = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      return nullptr;
    }
    compile_msm(g, it, n, rows);
  } else if (kind == CUBBS_HIP_BATCH_G1_MSM_COMMON_POINTS) {
    auto* it = static_cast<const CuBbsBbsGpuG1MsmCommonPointsBatchItem*>(items);
    if (!common_items_ok(it, n)) {
      status = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      return nullptr;
    }
    compile_common(g, it, n, rows);
  } else {
    status = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    return nullptr;
  }
  auto b = stage(g);
  b->rows = std::move(rows);
  b->n_total = b->rows.size();
  status = CUBBS_BBS_GPU_SUCCESS;
  return b;
}

void collect_primitive(Batch& b, CuBbsBbsGpuResult* results) {
  Ran r = fetch(b);
  b.h_out.assign(48 * b.rows.size(), 0);
  for (size_t k = 0; k < b.rows.size(); k++) {
    CuBbsBbsGpuByteBuffer buf{b.h_out.data() + 48 * k, 48, 0};
    row_result(b.rows[k], r.status.data(), r.out.data(), results[k], &buf);
  }
}

// ------------------------------------------------------------------ coalescing of small calls
// Many host threads issuing small independent calls (the Rust engine's
// per-item pairing checks and singleton MSMs from its worker threads) each
// pay a whole launch chain.  Calls of at most COALESCE_MAX_CALL_ITEMS items
// queue here; a caller that finds fewer than `inflight` combined calls
// running takes every queued call (up to COALESCE_MAX_ITEMS items) and runs
// them as one call, the others wait for their results.  Nothing waits for a
// timer: while combined calls run, new calls accumulate, so an idle library
// adds no latency.  This is the reference's pairing microbatcher
// (native/cuda/src/g1_msm_blst.cpp:1278-1406) without its gather window,
// applied to pairing_eq and item MSMs; results equal the direct call's (the
// items are independent rows).  CUBBS_HIP_COALESCE=0 turns it off.
constexpr size_t COALESCE_MAX_CALL_ITEMS = 16, COALESCE_MAX_ITEMS = 1024;

template <class Item>
struct Coalescer {
  struct Call {
    Item* items;
    size_t n;
    CuBbsBbsGpuResult* results;
    int32_t rc = CUBBS_BBS_GPU_ERR_INTERNAL;
    bool done = false;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Call*> pending;
  uint32_t running = 0;
  uint64_t combined_calls = 0, coalesced_calls = 0;

  template <class Direct>
  int32_t submit(Item* items, size_t n, CuBbsBbsGpuResult* results, Direct direct) {
    Call me{items, n, results};
    std::unique_lock<std::mutex> lk(mu);
    pending.push_back(&me);
    for (;;) {
      if (me.done) return me.rc;
      if (running < knobs().coalesce_inflight && !pending.empty()) {
        std::vector<Call*> take;
        size_t total = 0;
        while (!pending.empty() && (take.empty() || total + pending.front()->n <= COALESCE_MAX_ITEMS)) {
          take.push_back(pending.front());
          total += pending.front()->n;
          pending.pop_front();
        }
        running++;
        combined_calls++;
        coalesced_calls += take.size();
        lk.unlock();
        int32_t rc = CUBBS_BBS_GPU_ERR_INTERNAL;
        std::vector<Item> all;
        std::vector<CuBbsBbsGpuResult> res;
        try {
          all.reserve(total);
          for (Call* c : take) all.insert(all.end(), c->items, c->items + c->n);
          res.resize(total);
          rc = direct(all.data(), all.size(), res.data());
        } catch (...) {  // host memory exhausted
          rc = CUBBS_BBS_GPU_ERR_INTERNAL;
        }
        lk.lock();
        size_t off = 0;
        for (Call* c : take) {
          if (rc == CUBBS_BBS_GPU_SUCCESS)
            for (size_t i = 0; i < c->n; i++) {
              c->results[i] = res[off + i];
              c->items[i] = all[off + i];  // outputs written through the items (lengths)
            }
          c->rc = rc;
          c->done = true;
          off += c->n;
        }
        running--;
        cv.notify_all();
        continue;
      }
      cv.wait(lk);
    }
  }
};

Coalescer<CuBbsBbsGpuPairingEqItem>& pairing_coalescer() {
  static auto* c = new Coalescer<CuBbsBbsGpuPairingEqItem>();
  return *c;
}
Coalescer<CuBbsBbsGpuG1MsmItem>& msm_coalescer() {
  static auto* c = new Coalescer<CuBbsBbsGpuG1MsmItem>();
  return *c;
}

int32_t msm_direct(CuBbsBbsGpuG1MsmItem* items, size_t item_count, CuBbsBbsGpuResult* results) {
  return guard([&] {
    Prog g;
    std::vector<PrimRow> rows;
    compile_msm(g, items, item_count, rows);
    Ran r = run(g);
    for (size_t i = 0; i < item_count; i++)
      row_result(rows[i], r.status.data(), r.out.data(), results[i], &items[i].point_out);
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

int32_t pairing_direct(CuBbsBbsGpuPairingEqItem* items, size_t item_count, CuBbsBbsGpuResult* results) {
  return guard([&] {
    Prog g;
    std::vector<uint32_t> st(item_count);
    std::vector<bool> shape(item_count);
    for (size_t i = 0; i < item_count; i++) {
      const auto& it = items[i];
      st[i] = g.slot();
      shape[i] = fixed(it.lhs_g1, 48) && fixed(it.rhs_g1, 48) && fixed(it.lhs_g2, 96) &&
                 fixed(it.rhs_g2, 96);
      if (!shape[i]) continue;
      const uint32_t a = g.point(it.lhs_g1.data, st[i]);
      const uint32_t b = g.point(it.rhs_g1.data, st[i]);
      const uint32_t qa = (uint32_t)g.g2.size();
      g.g2.push_back(it.lhs_g2.data);
      g.g2.push_back(it.rhs_g2.data);
      g.p.pairs.push_back(PairJob{st[i], a, b, qa, qa + 1, 1u});
    }
    Ran ran = run(g);
    for (size_t i = 0; i < item_count; i++) {
      if (!shape[i] || ran.status[st[i]])
        set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
      else
        set(results[i], CUBBS_BBS_GPU_SUCCESS, ran.valid[st[i]] ? 1 : 0, 0);
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

}  // namespace cubbs

using namespace cubbs;

extern "C" {

int32_t cubbs_hip_coalesce_stats(uint64_t* pairing_combined, uint64_t* pairing_calls,
                                 uint64_t* msm_combined, uint64_t* msm_calls) {
  auto& p = pairing_coalescer();
  auto& m = msm_coalescer();
  {
    std::lock_guard<std::mutex> g(p.mu);
    if (pairing_combined) *pairing_combined = p.combined_calls;
    if (pairing_calls) *pairing_calls = p.coalesced_calls;
  }
  std::lock_guard<std::mutex> g(m.mu);
  if (msm_combined) *msm_combined = m.combined_calls;
  if (msm_calls) *msm_calls = m.coalesced_calls;
  return CUBBS_BBS_GPU_SUCCESS;
}

// reference header :264-267; semantics native/cuda/src/cubbs_bbs_cuda.cu:875-918
int32_t cubbs_bbs_gpu_g1_msm_batch_v2(uint32_t abi_version, CuBbsBbsGpuG1MsmItem* items,
                                      size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (!msm_items_ok(items, item_count)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t i = 0; i < item_count; i++)
    if (!outcap(items[i].point_out, 48)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  if (knobs().coalesce && item_count <= COALESCE_MAX_CALL_ITEMS)
    return msm_coalescer().submit(items, item_count, results, msm_direct);
  return msm_direct(items, item_count, results);
}

// reference header :269-275; semantics cubbs_bbs_cuda.cu:756-807 (validation),
// :920-1034.  Both entries compute the same matrix: output row r of group g is
// sum_t scalar(r, t) * points[t], scalars from scalar_rows[r * terms + t] or,
// when scalar_rows_be is set, from scalar_rows_be + r * stride + 32 t.
static int32_t common_points_entry(uint32_t abi_version,
                                   CuBbsBbsGpuG1MsmCommonPointsBatchItem* items,
                                   size_t group_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (group_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (!common_items_ok(items, group_count)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t gi = 0; gi < group_count; gi++) {
    if (items[gi].output_count && !items[gi].point_outs) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t r = 0; r < items[gi].output_count; r++)
      if (!outcap(items[gi].point_outs[r], 48)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  if (group_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    std::vector<PrimRow> rows;
    compile_common(g, items, group_count, rows);
    Ran ran = run(g);
    size_t k = 0;
    for (size_t gi = 0; gi < group_count; gi++)
      for (size_t r = 0; r < items[gi].output_count; r++, k++)
        row_result(rows[k], ran.status.data(), ran.out.data(), results[k], &items[gi].point_outs[r]);
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

int32_t cubbs_bbs_gpu_g1_msm_common_points_batch_v2(uint32_t abi_version,
                                                    CuBbsBbsGpuG1MsmCommonPointsBatchItem* items,
                                                    size_t group_count, CuBbsBbsGpuResult* results) {
  return common_points_entry(abi_version, items, group_count, results);
}

int32_t cubbs_bbs_gpu_g1_msm_fused_common_points_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1MsmCommonPointsBatchItem* items, size_t group_count,
    CuBbsBbsGpuResult* results) {
  return common_points_entry(abi_version, items, group_count, results);
}

// reference header :277-280; semantics cubbs_bbs_cuda.cu:1036-1078: row r is
// sum_t scalar_rows[r * terms + t] * points[t] + variable_scalars[r] * variable_points[r].
int32_t cubbs_bbs_gpu_g1_msm_common_points_plus_variable_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1MsmCommonPointsPlusVariableItem* items, size_t group_count,
    CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (group_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t gi = 0; gi < group_count; gi++) {
    const auto& it = items[gi];
    if ((it.term_count && (!it.points || !it.scalar_rows)) ||
        (it.output_count && (!it.variable_points || !it.variable_scalars || !it.point_outs)) ||
        it.term_count > SIZE_MAX / 32u ||
        (it.term_count && it.output_count > SIZE_MAX / it.term_count))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t t = 0; t < it.term_count; t++)
      if (!fixed(it.points[t], 48)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t r = 0; r < it.output_count; r++) {
      if (!outcap(it.point_outs[r], 48) || !fixed(it.variable_points[r], 48) ||
          !fixed(it.variable_scalars[r], 32))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      for (size_t t = 0; t < it.term_count; t++)
        if (!fixed(it.scalar_rows[r * it.term_count + t], 32)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    }
  }
  if (group_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    struct Row {
      uint32_t st, gst;
      size_t off;
    };
    std::vector<Row> rows;
    for (size_t gi = 0; gi < group_count; gi++) {
      const auto& it = items[gi];
      const uint32_t gst = g.slot();
      std::vector<uint32_t> gen, pts;
      if (!as_generators(it.points, it.term_count, gen)) {
        gen.clear();
        for (size_t t = 0; t < it.term_count; t++) pts.push_back(g.point(it.points[t].data, gst));
      }
      for (size_t r = 0; r < it.output_count; r++) {
        const uint32_t st = g.slot();
        const uint32_t first = g.p.n_partials;
        auto sc = [&](size_t t) { return it.scalar_rows[r * it.term_count + t].data; };
        common_row(g, pts, gen, it.term_count, st, sc);
        g.product(g.point(it.variable_points[r].data, st), g.scalar(it.variable_scalars[r].data, st));
        rows.push_back(Row{st, gst, g.compress(g.finish(first))});
      }
    }
    Ran ran = run(g);
    size_t k = 0;
    for (size_t gi = 0; gi < group_count; gi++) {
      for (size_t r = 0; r < items[gi].output_count; r++, k++) {
        const Row& row = rows[k];
        if (ran.status[row.st] || ran.status[row.gst]) {
          set(results[k], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
          continue;
        }
        std::memcpy(items[gi].point_outs[r].data, ran.out.data() + row.off, 48);
        items[gi].point_outs[r].len = 48;
        set(results[k], CUBBS_BBS_GPU_SUCCESS, ROUTE_COMMON, 48);
      }
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

// reference header :156-170, :282-284; semantics native/cuda/src/kernels/
// proof_verify.cuh:3-441: T1 = c Bbar + e^ Abar + r1^ D and
// T2 = sum_t common_scalar_rows[r * terms + t] common_points[t] + r3^ D, with
// proof_points [Bbar, Abar, D] and proof_scalars [c, e^, r1^, r3^] per row.
int32_t cubbs_bbs_gpu_proof_verify_msm_pair_batch_v2(uint32_t abi_version,
                                                     CuBbsBbsGpuProofVerifyMsmPairItem* items,
                                                     size_t group_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (group_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t gi = 0; gi < group_count; gi++) {
    const auto& it = items[gi];
    if (it.output_count == 0) continue;
    if (!it.common_term_count || !it.common_points || !it.common_scalar_rows || !it.proof_points ||
        !it.proof_scalars || !it.t1_outs || !it.t2_outs ||
        it.output_count > SIZE_MAX / it.common_term_count)
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t r = 0; r < it.output_count; r++)
      if (!outcap(it.t1_outs[r], 48) || !outcap(it.t2_outs[r], 48))
        return CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT;
    for (size_t t = 0; t < it.common_term_count; t++)
      if (!fixed(it.common_points[t], 48)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t r = 0; r < it.output_count; r++) {
      for (size_t t = 0; t < it.common_term_count; t++)
        if (!fixed(it.common_scalar_rows[r * it.common_term_count + t], 32))
          return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      for (size_t q = 0; q < 3; q++)
        if (!fixed(it.proof_points[r * 3 + q], 48)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      for (size_t q = 0; q < 4; q++)
        if (!fixed(it.proof_scalars[r * 4 + q], 32)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    }
  }
  if (group_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    struct Row {
      uint32_t st, gst;
      size_t t1, t2;
    };
    std::vector<Row> rows;
    for (size_t gi = 0; gi < group_count; gi++) {
      const auto& it = items[gi];
      if (it.output_count == 0) continue;
      const uint32_t gst = g.slot();
      std::vector<uint32_t> gen, pts;
      if (!as_generators(it.common_points, it.common_term_count, gen)) {
        gen.clear();
        for (size_t t = 0; t < it.common_term_count; t++)
          pts.push_back(g.point(it.common_points[t].data, gst));
      }
      for (size_t r = 0; r < it.output_count; r++) {
        const uint32_t st = g.slot();
        uint32_t pp[3], ps[4];
        for (int q = 0; q < 3; q++) pp[q] = g.point(it.proof_points[r * 3 + q].data, st);
        for (int q = 0; q < 4; q++) ps[q] = g.scalar(it.proof_scalars[r * 4 + q].data, st);
        uint32_t first = g.p.n_partials;
        g.product(pp[0], ps[0]);
        g.product(pp[1], ps[1]);
        g.product(pp[2], ps[2]);
        const size_t t1 = g.compress(g.finish(first));
        first = g.p.n_partials;
        auto sc = [&](size_t t) { return it.common_scalar_rows[r * it.common_term_count + t].data; };
        common_row(g, pts, gen, it.common_term_count, st, sc);
        g.product(pp[2], ps[3]);
        const size_t t2 = g.compress(g.finish(first));
        rows.push_back(Row{st, gst, t1, t2});
      }
    }
    Ran ran = run(g);
    size_t k = 0;
    for (size_t gi = 0; gi < group_count; gi++) {
      for (size_t r = 0; r < items[gi].output_count; r++, k++) {
        const Row& row = rows[k];
        if (ran.status[row.st] || ran.status[row.gst]) {
          set(results[k], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
          continue;
        }
        std::memcpy(items[gi].t1_outs[r].data, ran.out.data() + row.t1, 48);
        std::memcpy(items[gi].t2_outs[r].data, ran.out.data() + row.t2, 48);
        items[gi].t1_outs[r].len = 48;
        items[gi].t2_outs[r].len = 48;
        set(results[k], CUBBS_BBS_GPU_SUCCESS, ROUTE_COMMON, 96);
      }
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

// reference header :172-189, :286-292; semantics cubbs_bbs_cuda.cu:1106-1167:
// lhs = sum_t lhs_scalars[t] points[t], rhs = sum_t rhs_scalars[t] points[t];
// the lanes form splits the terms into lane_count contiguous ranges (range j =
// [j T / lanes, (j + 1) T / lanes)) and writes one lhs/rhs pair per range.
static int32_t dual_entry(size_t item_count, CuBbsBbsGpuResult* results,
                          const std::function<void(size_t, const CuBbsBbsGpuByteSlice*&,
                                                   const CuBbsBbsGpuByteSlice*&,
                                                   const CuBbsBbsGpuByteSlice*&, size_t&, size_t&,
                                                   uint8_t*&, uint8_t*&)>& get,
                          const std::function<void(size_t, size_t)>& done) {
  return guard([&] {
    Prog g;
    struct Row {
      uint32_t st;
      std::vector<size_t> lo, ro;
    };
    std::vector<Row> rows(item_count);
    for (size_t i = 0; i < item_count; i++) {
      const CuBbsBbsGpuByteSlice *pts, *ls, *rs;
      size_t terms, lanes;
      uint8_t *lo, *ro;
      get(i, pts, ls, rs, terms, lanes, lo, ro);
      Row& row = rows[i];
      row.st = g.slot();
      std::vector<uint32_t> pt(terms);
      for (size_t t = 0; t < terms; t++) pt[t] = g.point(pts[t].data, row.st);
      for (size_t j = 0; j < lanes; j++) {
        const size_t b = j * terms / lanes, e = (j + 1) * terms / lanes;
        uint32_t first = g.p.n_partials;
        for (size_t t = b; t < e; t++) g.product(pt[t], g.scalar(ls[t].data, row.st));
        row.lo.push_back(g.compress(g.finish(first)));
        first = g.p.n_partials;
        for (size_t t = b; t < e; t++) g.product(pt[t], g.scalar(rs[t].data, row.st));
        row.ro.push_back(g.compress(g.finish(first)));
      }
    }
    Ran ran = run(g);
    for (size_t i = 0; i < item_count; i++) {
      const CuBbsBbsGpuByteSlice *pts, *ls, *rs;
      size_t terms, lanes;
      uint8_t *lo, *ro;
      get(i, pts, ls, rs, terms, lanes, lo, ro);
      if (ran.status[rows[i].st]) {
        set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
        continue;
      }
      for (size_t j = 0; j < lanes; j++) {
        std::memcpy(lo + 48 * j, ran.out.data() + rows[i].lo[j], 48);
        std::memcpy(ro + 48 * j, ran.out.data() + rows[i].ro[j], 48);
      }
      done(i, 48 * lanes);
      set(results[i], CUBBS_BBS_GPU_SUCCESS, ROUTE_DUAL, 96 * lanes);
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

int32_t cubbs_bbs_gpu_g1_dual_msm_same_points_batch_v2(uint32_t abi_version,
                                                       CuBbsBbsGpuG1DualMsmSamePointsItem* items,
                                                       size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (!it.term_count || !it.points || !it.lhs_scalars || !it.rhs_scalars ||
        !outcap(it.lhs_out, 48) || !outcap(it.rhs_out, 48))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t t = 0; t < it.term_count; t++)
      if (!fixed(it.points[t], 48) || !fixed(it.lhs_scalars[t], 32) || !fixed(it.rhs_scalars[t], 32))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return dual_entry(
      item_count, results,
      [&](size_t i, const CuBbsBbsGpuByteSlice*& pts, const CuBbsBbsGpuByteSlice*& ls,
          const CuBbsBbsGpuByteSlice*& rs, size_t& terms, size_t& lanes, uint8_t*& lo,
          uint8_t*& ro) {
        pts = items[i].points;
        ls = items[i].lhs_scalars;
        rs = items[i].rhs_scalars;
        terms = items[i].term_count;
        lanes = 1;
        lo = items[i].lhs_out.data;
        ro = items[i].rhs_out.data;
      },
      [&](size_t i, size_t bytes) {
        items[i].lhs_out.len = bytes;
        items[i].rhs_out.len = bytes;
      });
}

int32_t cubbs_bbs_gpu_g1_dual_msm_same_points_lanes_batch_v2(
    uint32_t abi_version, CuBbsBbsGpuG1DualMsmSamePointsLanesItem* items, size_t item_count,
    CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (it.lane_count > SIZE_MAX / 96u) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    const size_t bytes = it.lane_count * 48u;
    if (!it.term_count || !it.lane_count || it.lane_count > it.term_count || !it.points ||
        !it.lhs_scalars || !it.rhs_scalars || !outcap(it.lhs_outs, bytes) ||
        !outcap(it.rhs_outs, bytes))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t t = 0; t < it.term_count; t++)
      if (!fixed(it.points[t], 48) || !fixed(it.lhs_scalars[t], 32) || !fixed(it.rhs_scalars[t], 32))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return dual_entry(
      item_count, results,
      [&](size_t i, const CuBbsBbsGpuByteSlice*& pts, const CuBbsBbsGpuByteSlice*& ls,
          const CuBbsBbsGpuByteSlice*& rs, size_t& terms, size_t& lanes, uint8_t*& lo,
          uint8_t*& ro) {
        pts = items[i].points;
        ls = items[i].lhs_scalars;
        rs = items[i].rhs_scalars;
        terms = items[i].term_count;
        lanes = items[i].lane_count;
        lo = items[i].lhs_outs.data;
        ro = items[i].rhs_outs.data;
      },
      [&](size_t i, size_t bytes) {
        items[i].lhs_outs.len = bytes;
        items[i].rhs_outs.len = bytes;
      });
}

// reference header :191-195, :294-297; semantics native/cuda/src/g1_msm_blst.cpp:349-371,
// 1011-1056 (per-item error codes, lengths checked per item).
int32_t cubbs_bbs_gpu_g1_add_batch_v2(uint32_t abi_version, CuBbsBbsGpuG1AddItem* items,
                                      size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    std::vector<uint32_t> st(item_count);
    std::vector<size_t> off(item_count);
    std::vector<bool> shape(item_count);
    const uint8_t one[32] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                             0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1};
    for (size_t i = 0; i < item_count; i++) {
      const auto& it = items[i];
      shape[i] = fixed(it.lhs, 48) && fixed(it.rhs, 48) && outcap(it.point_out, 48);
      st[i] = g.slot();
      if (!shape[i]) continue;
      const uint32_t first = g.p.n_partials;
      const uint32_t sc = g.scalar(one, st[i]);
      g.product(g.point(it.lhs.data, st[i]), sc);
      g.product(g.point(it.rhs.data, st[i]), sc);
      off[i] = g.compress(g.finish(first));
    }
    Ran ran = run(g);
    for (size_t i = 0; i < item_count; i++) {
      if (!shape[i] || ran.status[st[i]]) {
        set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
        continue;
      }
      std::memcpy(items[i].point_out.data, ran.out.data() + off[i], 48);
      items[i].point_out.len = 48;
      set(results[i], CUBBS_BBS_GPU_SUCCESS, 0, 48);
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

// reference header :197-202, :299-301; semantics g1_msm_blst.cpp:373-404:
// value = [e(lhs_g1, lhs_g2) == e(rhs_g1, rhs_g2)], checked as
// e(lhs_g1, lhs_g2) e(-rhs_g1, rhs_g2) == 1; undecodable or non-subgroup
// inputs give code 1 for the item.
int32_t cubbs_bbs_gpu_pairing_eq_batch_v2(uint32_t abi_version, const CuBbsBbsGpuPairingEqItem* items,
                                          size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  // the items are only read: a private copy serves both paths
  std::vector<CuBbsBbsGpuPairingEqItem> own;
  try {
    own.assign(items, items + item_count);
  } catch (...) {
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  if (knobs().coalesce && item_count <= COALESCE_MAX_CALL_ITEMS)
    return pairing_coalescer().submit(own.data(), item_count, results, pairing_direct);
  return pairing_direct(own.data(), item_count, results);
}

// reference header :204-210, :302-304; semantics g1_msm_blst.cpp:1422-1640: one
// result for the whole set.  With common G2 points on each side the equation
// e(sum c_i lhs_i, Q_l) == e(sum c_i rhs_i, Q_r) is checked once (coefficient
// MSMs on the device, then one pairing check); otherwise every item's
// e(lhs_i, Ql_i) == e(rhs_i, Qr_i) must hold.
int32_t cubbs_bbs_gpu_pairing_eq_aggregate_v2(uint32_t abi_version,
                                              const CuBbsBbsGpuPairingEqScaledItem* items,
                                              size_t item_count, CuBbsBbsGpuResult* result) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && !items) || !result)
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  set(*result, CUBBS_BBS_GPU_SUCCESS, 1, 0);
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (!fixed(it.lhs_g1, 48) || !fixed(it.rhs_g1, 48) || !fixed(it.lhs_g2, 96) ||
        !fixed(it.rhs_g2, 96) || !fixed(it.coefficient, 32)) {
      set(*result, CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
      return CUBBS_BBS_GPU_SUCCESS;
    }
  }
  bool common = true;
  for (size_t i = 1; i < item_count && common; i++)
    common = std::memcmp(items[i].lhs_g2.data, items[0].lhs_g2.data, 96) == 0 &&
             std::memcmp(items[i].rhs_g2.data, items[0].rhs_g2.data, 96) == 0;
  if (!common) {
    std::vector<CuBbsBbsGpuPairingEqItem> plain(item_count);
    std::vector<CuBbsBbsGpuResult> res(item_count);
    for (size_t i = 0; i < item_count; i++)
      plain[i] = CuBbsBbsGpuPairingEqItem{items[i].lhs_g1, items[i].lhs_g2, items[i].rhs_g1,
                                          items[i].rhs_g2};
    const int32_t st = cubbs_bbs_gpu_pairing_eq_batch_v2(abi_version, plain.data(), item_count, res.data());
    if (st != CUBBS_BBS_GPU_SUCCESS) {
      set(*result, st, 0, 0);
      return CUBBS_BBS_GPU_SUCCESS;
    }
    for (const auto& r : res) {
      if (r.code != CUBBS_BBS_GPU_SUCCESS) {
        set(*result, r.code, 0, 0);
        return CUBBS_BBS_GPU_SUCCESS;
      }
      if (r.value != 1) {
        set(*result, CUBBS_BBS_GPU_SUCCESS, 0, 0);
        return CUBBS_BBS_GPU_SUCCESS;
      }
    }
    return CUBBS_BBS_GPU_SUCCESS;
  }
  const int32_t rc = guard([&] {
    Prog g;
    const uint32_t st = g.slot();
    uint32_t first = g.p.n_partials;
    std::vector<uint32_t> sc(item_count);
    for (size_t i = 0; i < item_count; i++) sc[i] = g.scalar(items[i].coefficient.data, st);
    for (size_t i = 0; i < item_count; i++) g.product(g.point(items[i].lhs_g1.data, st), sc[i]);
    const uint32_t a = g.finish(first);
    first = g.p.n_partials;
    for (size_t i = 0; i < item_count; i++) g.product(g.point(items[i].rhs_g1.data, st), sc[i]);
    const uint32_t b = g.finish(first);
    g.g2.push_back(items[0].lhs_g2.data);
    g.g2.push_back(items[0].rhs_g2.data);
    g.p.pairs.push_back(PairJob{st, a, b, 0, 1, 1u});
    Ran ran = run(g);
    if (ran.status[st])
      set(*result, CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
    else
      set(*result, CUBBS_BBS_GPU_SUCCESS, ran.valid[st] ? 1 : 0, 0);
    return CUBBS_BBS_GPU_SUCCESS;
  });
  if (rc != CUBBS_BBS_GPU_SUCCESS) set(*result, rc, 0, 0);
  return CUBBS_BBS_GPU_SUCCESS;
}

// reference header :121-127, :259-262; validation native/cuda/src/cubbs_bbs_cuda.cu:837-873
// (the reference body is a stub).  Computes B = P1 + Q1 domain + sum_i H_i m_i
// (bbsplus/signature.rs:620-640) with domain = calculate_domain(PK, header, L)
// (utils/util.rs:147), message scalars given; b_out receives compressed B.
int32_t cubbs_bbs_gpu_prepare_b_batch_v2(uint32_t abi_version, CuBbsBbsGpuPrepareBItem* items,
                                         size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    if (!fixed(it.public_key, CUBBS_BBS_PUBLIC_KEY_LEN) || !fixed(it.header, CUBBS_BBS_2023_HEADER_LEN) ||
        (it.message_count && !it.message_scalars) || !outcap(it.b_out, 48) ||
        it.message_count >= (1u << 20))
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    for (size_t j = 0; j < it.message_count; j++)
      if (!fixed(it.message_scalars[j], 32)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    std::vector<uint32_t> st(item_count);
    std::vector<size_t> off(item_count);
    for (size_t i = 0; i < item_count; i++) {
      const auto& it = items[i];
      st[i] = g.slot();
      const uint32_t L = (uint32_t)it.message_count;
      const uint32_t sc = g.p.scalars(L + 2);
      g.p.fills.push_back(TermFill{sc, 0, 0, 2});  // P1 scalar 1
      const uint32_t pk = (uint32_t)g.g2.size();
      g.g2.push_back(it.public_key.data);
      // pk_slot is an index into g.g2 until stage() maps it
      g.p.dom.push_back(DomainJob{g.p.bytes(it.header.data, it.header.len), (uint32_t)it.header.len,
                                  pk, L, sc + 1});
      for (uint32_t m = 0; m < L; m++) g.scalar_at(sc + 2 + m, it.message_scalars[m].data, st[i], 0);
      std::vector<uint32_t> gen(L + 2);
      for (uint32_t t = 0; t < L + 2; t++) gen[t] = t;
      const uint32_t first = g.p.n_partials;
      g.comb(sc, gen);
      off[i] = g.compress(g.finish(first));
    }
    Ran ran = run(g);
    for (size_t i = 0; i < item_count; i++) {
      if (ran.status[st[i]]) {
        set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
        continue;
      }
      std::memcpy(items[i].b_out.data, ran.out.data() + off[i], 48);
      items[i].b_out.len = 48;
      set(results[i], CUBBS_BBS_GPU_SUCCESS, ROUTE_COMMON, 48);
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

// reference header :212-225, :306-308; semantics native/cuda/src/kernels/
// proof_gen.cuh:17-110,192-447: with bases B (compressed_b) and A (signature_a)
// and multipliers s0..s5 = [r2, r1 r2, r1 r2, r1 r2 e, r1 r2 e~, r2 r1~]:
// d = s0 B, abar = s1 A, bbar = s2 B - s3 A, t1 = s4 A + s5 B (s6 = r2 r3~ is
// not used on the device).  Multipliers are reduced mod r.
int32_t cubbs_bbs_gpu_proof_gen_t2_prepare_batch_v2(uint32_t abi_version,
                                                    CuBbsBbsGpuProofGenT2PrepareItem* items,
                                                    size_t item_count, CuBbsBbsGpuResult* results) {
  if (abi_version != CUBBS_BBS_GPU_ABI_VERSION || (item_count && (!items || !results)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  for (size_t i = 0; i < item_count; i++) {
    const auto& it = items[i];
    bool ok = fixed(it.compressed_b, 48) && fixed(it.signature_a, 48) && outcap(it.d_out, 48) &&
              outcap(it.abar_out, 48) && outcap(it.bbar_out, 48) && outcap(it.t1_out, 48);
    for (size_t s = 0; s < 7 && ok; s++) ok = fixed(it.scalars_be[s], 32);
    if (!ok) {
      set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    }
  }
  if (item_count == 0) return CUBBS_BBS_GPU_SUCCESS;
  return guard([&] {
    Prog g;
    std::vector<uint32_t> st(item_count);
    std::vector<size_t> off(item_count * 4);
    for (size_t i = 0; i < item_count; i++) {
      const auto& it = items[i];
      st[i] = g.slot();
      const uint32_t B = g.point(it.compressed_b.data, st[i]);
      const uint32_t A = g.point(it.signature_a.data, st[i]);
      uint32_t s[6];
      for (int k = 0; k < 6; k++) s[k] = g.scalar(it.scalars_be[k].data, st[i], k == 3 ? 3u : 2u);
      uint32_t first = g.p.n_partials;
      g.product(B, s[0]);
      off[4 * i + 0] = g.compress(g.finish(first));
      first = g.p.n_partials;
      g.product(A, s[1]);
      off[4 * i + 1] = g.compress(g.finish(first));
      first = g.p.n_partials;
      g.product(B, s[2]);
      g.product(A, s[3]);  // s3 stored negated
      off[4 * i + 2] = g.compress(g.finish(first));
      first = g.p.n_partials;
      g.product(A, s[4]);
      g.product(B, s[5]);
      off[4 * i + 3] = g.compress(g.finish(first));
    }
    Ran ran = run(g);
    for (size_t i = 0; i < item_count; i++) {
      if (ran.status[st[i]]) {
        set(results[i], CUBBS_BBS_GPU_ERR_INVALID_INPUT, 0, 0);
        continue;
      }
      CuBbsBbsGpuByteBuffer* outs[4] = {&items[i].d_out, &items[i].abar_out, &items[i].bbar_out,
                                        &items[i].t1_out};
      for (int k = 0; k < 4; k++) {
        std::memcpy(outs[k]->data, ran.out.data() + off[4 * i + k], 48);
        outs[k]->len = 48;
      }
      set(results[i], CUBBS_BBS_GPU_SUCCESS, ROUTE_ITEM, 48);
    }
    return CUBBS_BBS_GPU_SUCCESS;
  });
}

}  // extern "C"
