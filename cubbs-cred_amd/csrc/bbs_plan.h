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
  std::vector<MsgHashJob> msg;     // message hashes (MAP DST), inputs laid out below
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
  // groups, and how many of pairs[] (at its front) are combined checks: the
  // supergroup checks first (n_super_pairs), then the group checks
  std::vector<VarJob> agg_var;       // run in the same launch as `var`, after it
  std::vector<VarSum> var_sum;       // multi-term products (same launch, first)
  uint32_t n_agg_slots = 0, n_group_pairs = 0, n_super_pairs = 0;
  // (x, y) scalar index pairs of the summed term fills (TermFill mode 3)
  std::vector<uint32_t> sum_pairs;
  // work that runs only for the items of failing groups, after the group
  // checks: comb chunks (each gated by its group's status slot) and the
  // sums that form the items' own check points
  std::vector<CombChunk> fb_chunks;
  std::vector<uint32_t> fb_gate;
  std::vector<GatedReduce> fb_red;
  std::vector<VarSum> fb_sum;      // gated multi-term products (their own launch)
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
  size_t tail_off = 0;            // b_1 / b_2 input words of the message hashes
  uint32_t tail_blocks = 0;
  // Lay out `n` bytes at a 16-byte aligned offset as big-endian 32-bit words
  // (the byte order of SHA-256's words, so the device loads them as is).
  static void be_words(uint8_t* q, size_t n) {
    for (size_t i = 0; i + 4 <= n; i += 4) {
      uint32_t w;
      std::memcpy(&w, q + i, 4);
      w = __builtin_bswap32(w);
      std::memcpy(q + i, &w, 4);
    }
  }
  size_t aligned16(size_t n) {
    const size_t off = (blob.size() + 15) & ~(size_t)15;
    blob.resize(off + n, 0);
    return off;
  }
  // hash_to_scalar(m, map_dst) of one message into scalars[out_index]: the
  // input of expand_message_xmd's b_0 after Z_pad, m || I2OSP(48, 2) || 0x00 ||
  // DST || I2OSP(|DST|, 1), with SHA-256's padding (RFC 9380 5.3.1, FIPS
  // 180-4 5.1.1), in whole blocks; the hashing itself runs on the device.
  void hash_message(const uint8_t* m, size_t len, uint32_t out_index) {
    const size_t dl = map_dst.size();
    const size_t bytes_in = len + 3 + dl + 1;
    const size_t nb = (bytes_in + 9 + 63) / 64;
    const size_t off = aligned16(64 * nb);
    uint8_t* q = blob.data() + off;
    if (len) std::memcpy(q, m, len);
    q[len] = 0;
    q[len + 1] = 48;
    q[len + 2] = 0;
    if (dl) std::memcpy(q + len + 3, map_dst.data(), dl);
    q[len + 3 + dl] = (uint8_t)dl;
    q[bytes_in] = 0x80;
    const uint64_t bits = (uint64_t)(64 + bytes_in) * 8;  // the Z_pad block counts
    for (int k = 0; k < 8; k++) q[64 * nb - 1 - k] = (uint8_t)(bits >> (8 * k));
    be_words(q, 64 * nb);
    msg.push_back(MsgHashJob{off, (uint32_t)nb, out_index});
  }
  // The inputs of b_1 and b_2 for the message hashes: 32 bytes (filled on
  // the device) || I2OSP(i, 1) || DST || I2OSP(|DST|, 1), padded.
  void message_tails() {
    const size_t dl = map_dst.size();
    const size_t bytes_in = 32 + 1 + dl + 1;
    const size_t nb = (bytes_in + 9 + 63) / 64;
    tail_off = aligned16(2 * 64 * nb);
    tail_blocks = (uint32_t)nb;
    for (int i = 1; i <= 2; i++) {
      uint8_t* q = blob.data() + tail_off + 64 * nb * (i - 1);
      q[32] = (uint8_t)i;
      if (dl) std::memcpy(q + 33, map_dst.data(), dl);
      q[33 + dl] = (uint8_t)dl;
      q[bytes_in] = 0x80;
      const uint64_t bits = (uint64_t)bytes_in * 8;
      for (int k = 0; k < 8; k++) q[64 * nb - 1 - k] = (uint8_t)(bits >> (8 * k));
      be_words(q, 64 * nb);
    }
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
  // The same for the gated comb launch: chunks that run only when
  // valid[gate] == 0.  Returns the first of their consecutive partials.
  uint32_t comb_gated(uint32_t term, uint32_t count, uint32_t chunk, uint32_t gate) {
    const uint32_t first = n_partials;
    for (uint32_t t = 0; t < count; t += chunk) {
      uint32_t c = std::min(chunk, count - t);
      fb_chunks.push_back(CombChunk{term + t, c, partials(1)});
      fb_gate.push_back(gate);
    }
    return first;
  }
  // partials[partial] = sum of scalars[s] * points[p] over the (p, s) terms
  // given (at most VAR_TERMS_MAX), as one multi-term product job.
  void var_terms(const std::vector<std::pair<uint32_t, uint32_t>>& ps, uint32_t partial) {
    VarSum v{};
    v.count = (uint32_t)ps.size();
    for (uint32_t t = 0; t < v.count; t++) {
      v.point[t] = ps[t].first;
      v.scalar[t] = ps[t].second;
    }
    v.partial = partial;
    var_sum.push_back(v);
  }
  // One dual job over the points p[0..count): partials[partial] = sum of
  // scalars[s1[t]] * points[p[t]] and partials[partial2] = the same with
  // s2.  gate != PRIM_NO_SLOT: a gated job (fb_sum), run only when its
  // group's check did not hold.
  void var_dual(const uint32_t* p, const uint32_t* s1, const uint32_t* s2, uint32_t count,
                uint32_t partial, uint32_t partial2, uint32_t gate = PRIM_NO_SLOT) {
    VarSum v{};
    v.count = count;
    for (uint32_t t = 0; t < count; t++) {
      v.point[t] = p[t];
      v.scalar[t] = s1[t];
      v.scalar2[t] = s2[t];
    }
    v.partial = partial;
    v.partial2 = partial2;
    v.flags = VAR_SUM_DUAL;
    if (gate != PRIM_NO_SLOT) {
      v.flags |= VAR_SUM_GATED;
      v.gate = gate;
      fb_sum.push_back(v);
    } else {
      var_sum.push_back(v);
    }
  }
  // A summed term fill: scalars[dst] = sum of scalars[x] * scalars[y] over
  // the (x, y) pairs given.
  void term_sum(uint32_t dst, const std::vector<std::pair<uint32_t, uint32_t>>& xy) {
    const uint32_t first = (uint32_t)(sum_pairs.size() / 2);
    for (const auto& q : xy) {
      sum_pairs.push_back(q.first);
      sum_pairs.push_back(q.second);
    }
    fills.push_back(TermFill{dst, first, (uint32_t)xy.size(), 3});
  }
};

// Copy the plan into one pinned buffer + device mirror and allocate the work
// area (from the thread's Arena when one is set).
void materialize(Batch& b, Plan& p);
// Group the plain pairing checks e(a_i, PK) e(b_i, -G2) = 1 of p.pairs by key
// into groups of at most group_max (status slots from nv on) and add their
// random-linear-combination checks (the products c_i a_i, c_i b_i are
// planned here); see bbs_batch.cpp.
// super_max >= 2: consecutive groups of one key, up to super_max of them,
// are also checked together first (a supergroup), and their groups only when
// that check fails.
void aggregate_pairs(Plan& p, uint32_t nv, uint32_t group_max, uint32_t super_max = 0);
uint32_t agg_group_max();
uint32_t agg_super_max();
uint32_t agg_straus();
// Lock the context, make generators [0, max_gen_slots) and the given public
// keys resident, snapshot the resident pointers into the batch.
void attach(Batch& b, uint32_t max_gen_slots, const std::vector<const uint8_t*>& pks,
            std::vector<uint32_t>& pk_slot);

}  // namespace cubbs
