// Host-side plan of a device batch: job lists, the input blob and slot
// counters, turned into device arrays by materialize() (bbs_batch.cpp).
#pragma once
#include <algorithm>
#include <cstring>
#include <vector>

#include "bbs_batch.h"

namespace cubbs {

// ------------------------------------------------------------------ planning helpers
template <class T>
inline size_t put(std::vector<uint8_t>& blob, const T* p, size_t n) {
  size_t off = blob.size();
  blob.resize(off + n * sizeof(T));
  if (n) std::memcpy(blob.data() + off, p, n * sizeof(T));
  return off;
}

struct Plan {
  // host-side job lists
  std::vector<H2sJob> h2s;
  std::vector<DomainJob> dom;
  std::vector<DecodeG1Job> dec_g1;
  std::vector<DecodeFrJob> dec_fr;
  std::vector<TermFill> fills;
  std::vector<VarJob> var;
  std::vector<CombChunk> chunks;
  std::vector<uint32_t> term_gen;
  std::vector<ReduceGroup> groups1, groups2;
  std::vector<uint32_t> out1, out2;
  std::vector<PairJob> pairs;
  // partial-sum tree levels (primitive MSMs): level l sums groups of level-l
  // inputs into new partial slots; the last level feeds groups1
  std::vector<std::vector<ReduceGroup>> pst_groups;
  std::vector<std::vector<uint32_t>> pst_out;
  std::vector<CompressJob> comp;  // affine point -> 48-byte encoding in the output area
  std::vector<SignJob> sign;
  std::vector<PvChallengeJob> pvch;
  std::vector<PgJob> pg;
  std::vector<uint32_t> disc_index, disc_scalar, und_scalar;
  // aggregated pairing checks: products c_i a_i, c_i b_i of the grouped
  // checks (summed per group by groups2 / out2), extra status slots for the
  // groups, and how many of pairs[] (at its front) are group checks
  std::vector<VarJob> agg_var;       // run in the same launch as `var`, after it
  std::vector<PointCopy> pt_copy;    // affine inputs of the group sums (pre-scaled checks)
  uint32_t n_agg_slots = 0, n_group_pairs = 0;
  std::vector<PipTerm> pip_terms;  // terms of the Pippenger MSMs
  std::vector<uint32_t> pip_out;   // per Pippenger MSM: its output partial slot
  std::vector<uint8_t> blob;       // signatures, proofs, headers, messages, keys, seeds
  std::vector<uint8_t> map_dst;    // DST of message scalars
  uint32_t n_scalars = 0, n_points = 0, n_partials = 0;
  size_t out_bytes = 0;            // signature / proof output area
  bool seeded = false;
  size_t seeds_off = 0;

  uint32_t scalars(uint32_t k) { uint32_t s = n_scalars; n_scalars += k; return s; }
  uint32_t points(uint32_t k) { uint32_t s = n_points; n_points += k; return s; }
  uint32_t partials(uint32_t k) { uint32_t s = n_partials; n_partials += k; return s; }
  // A new Pippenger MSM (terms added to pip_terms with its index); its sum
  // lands in one new partial slot.
  uint32_t pip_job() {
    pip_out.push_back(partials(1));
    return (uint32_t)pip_out.size() - 1;
  }
  size_t bytes(const void* p, size_t n) {
    size_t off = blob.size();
    blob.resize(off + n);
    if (n) std::memcpy(blob.data() + off, p, n);
    return off;
  }
  // points[out] = sum of partials [begin, begin + count).  A long range is
  // first summed in parallel tree levels (runs of TREE_FANIN partials per lane
  // into new partial slots) down to at most FINAL_MAX partials.
  static constexpr uint32_t TREE_FANIN = 8, FINAL_MAX = 16;
  void reduce_to(uint32_t begin, uint32_t count, uint32_t out) {
    size_t level = 0;
    while (count > FINAL_MAX) {
      if (pst_groups.size() <= level) {
        pst_groups.emplace_back();
        pst_out.emplace_back();
      }
      const uint32_t groups = (count + TREE_FANIN - 1) / TREE_FANIN;
      const uint32_t first = partials(groups);
      for (uint32_t g = 0; g < groups; g++) {
        const uint32_t b = begin + g * TREE_FANIN;
        pst_groups[level].push_back(ReduceGroup{b, std::min(TREE_FANIN, begin + count - b)});
        pst_out[level].push_back(first + g);
      }
      begin = first;
      count = groups;
      level++;
    }
    groups1.push_back(ReduceGroup{begin, count});
    out1.push_back(out);
  }
  // Fixed-base terms [term, term + count) in chunks, partials appended to the
  // caller's running partial range.
  void comb(uint32_t term, uint32_t count, uint32_t chunk) {
    for (uint32_t t = 0; t < count; t += chunk) {
      uint32_t c = std::min(chunk, count - t);
      chunks.push_back(CombChunk{term + t, c, partials(1)});
    }
  }
};

// Copy the plan into one pinned buffer + device mirror and allocate the work
// area (from the thread's Arena when one is set).
void materialize(Batch& b, Plan& p);
// Group the plain pairing checks e(a_i, PK) e(b_i, -G2) = 1 of p.pairs by key
// into groups of at most group_max (status slots from nv on) and add their
// random-linear-combination checks; see bbs_batch.cpp.
// scaled: the checks' points are already multiplied by their coefficients
// (the planner drew them), so the group sums add the points themselves;
// otherwise the products c_i a_i, c_i b_i are planned here.
void aggregate_pairs(Plan& p, uint32_t nv, uint32_t group_max, bool scaled = false);
uint32_t agg_group_max();
// Lock the context, make generators [0, max_gen_slots) and the given public
// keys resident, snapshot the resident pointers into the batch.
void attach(Batch& b, uint32_t max_gen_slots, const std::vector<const uint8_t*>& pks,
            std::vector<uint32_t>& pk_slot);

}  // namespace cubbs
