// Batch planner: turns Sign / Verify / ProofGen / ProofVerify items into the
// flat job arrays of the device pipeline (bbs_pipeline.hip), stages them in
// one pinned upload, enqueues the stages on the batch's own stream and reads
// back per-item results.  A batch can be launched repeatedly with its inputs
// resident on the device (used by the benchmark and the serving loop).
//
// Semantics per operation follow the reference engine entry points
// (crates/cubbs-bbs/src/protocol.rs): sign :28-40, verify :318-326,
// proof_gen :423-437, proof_verify :772-793.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_batch.h"
#include "bbs_plan.h"
#include "bls12_381_consts.hpp"
#include "host_ctx.h"

namespace cubbs {

static bool slice_ok(const CuBbsBbsGpuByteSlice& s) { return s.len == 0 || s.data != nullptr; }

static const char* kApiId = "BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_";

// Terms per comb lane: the calling thread's forced shape (the tuner's
// candidate, or CUBBS_HIP_BATCH_COMB_CHUNK), else CUBBS_HIP_COMB_CHUNK, else
// (primitive MSM programs only, `tuned`) the comb tuner's choice for the
// batch's size class, else 4.  The whole-operation pipelines use
// pipeline_comb_chunk: their comb runs beside decode and the GLV products
// and feeds the reduction, which the comb-only tuning does not see.
uint32_t comb_chunk_terms(uint64_t batch_terms, bool tuned) {
  if (tune_force().comb_chunk) return tune_force().comb_chunk;
  if (knobs().comb_chunk) return knobs().comb_chunk;
  if (!tuned) return 4;
  const uint32_t t = tuned_comb_chunk(batch_terms);
  return t ? t : 4;
}
// As comb_chunk_terms for a Verify (kind 0) or ProofVerify (1) batch of
// `items`: the pipeline tuner's choice for its size class, else 4.
static uint32_t pipeline_comb_chunk(int kind, uint64_t items) {
  if (tune_force().comb_chunk) return tune_force().comb_chunk;
  if (knobs().comb_chunk) return knobs().comb_chunk;
  uint32_t chunk = 0, group = 0, super = 0;
  return tuned_pipeline(kind, items, chunk, group, super) && chunk ? chunk : 4;
}

// ------------------------------------------------------------------ arenas
static thread_local Arena* tl_arena = nullptr;
void set_thread_arena(Arena* a) { tl_arena = a; }
static thread_local bool tl_single_stream = false;
void set_thread_single_stream(bool on) { tl_single_stream = on; }

static uint8_t* arena_take(std::vector<Arena::Block>& blocks, size_t& held, size_t bytes,
                           size_t min_block, bool host) {
  bytes = align256(std::max<size_t>(bytes, 1));
  for (auto& bl : blocks) {
    if (bl.cap - bl.used >= bytes) {
      uint8_t* p = bl.p + bl.used;
      bl.used += bytes;
      return p;
    }
  }
  const size_t cap = std::max(bytes, min_block);
  uint8_t* p = nullptr;
  if (host)
    check(hipHostMalloc(&p, cap, hipHostMallocDefault));
  else
    check(hipMalloc(&p, cap));
  blocks.push_back(Arena::Block{p, cap, bytes});
  held += cap;
  return p;
}
uint8_t* Arena::alloc_dev(size_t bytes) { return arena_take(dev, dev_bytes, bytes, min_block, false); }
uint8_t* Arena::alloc_host(size_t bytes) {
  return arena_take(host, host_bytes, bytes, min_block / 4, true);
}
void Arena::reset() {
  for (auto& bl : dev) bl.used = 0;
  for (auto& bl : host) bl.used = 0;
}
Arena::~Arena() {
  for (hipStream_t s : streams)
    if (s) (void)hipStreamSynchronize(s);
  for (hipStream_t s : aux)
    if (s) (void)hipStreamSynchronize(s);
  for (auto& bl : dev) (void)hipFree(bl.p);
  for (auto& bl : host) (void)hipHostFree(bl.p);
  for (int i = 0; i < 4; i++) {  // streams belong to the context's pool
    if (fork[i]) (void)hipEventDestroy(fork[i]);
    if (join[i]) (void)hipEventDestroy(join[i]);
  }
}

// Copy the plan into one pinned buffer + device mirror; allocate the work area.
// Counting sort of job records by a small integer key (stable).
template <class T, class F>
static void sort_jobs(std::vector<T>& v, F key) {
  if (v.size() < 2) return;
  uint32_t kmax = 0;
  for (const T& x : v) kmax = std::max(kmax, key(x));
  std::vector<uint32_t> at((size_t)kmax + 2, 0);
  for (const T& x : v) at[key(x) + 1]++;
  for (size_t i = 1; i < at.size(); i++) at[i] += at[i - 1];
  std::vector<T> out(v.size());
  for (const T& x : v) out[at[key(x)]++] = x;
  v.swap(out);
}

// The same for reduce groups and their parallel output slots.
static void sort_groups(std::vector<ReduceGroup>& g, std::vector<uint32_t>& out) {
  if (g.size() < 2 || out.size() != g.size()) return;
  std::vector<std::pair<ReduceGroup, uint32_t>> z(g.size());
  for (size_t i = 0; i < g.size(); i++) z[i] = {g[i], out[i]};
  sort_jobs(z, [](const std::pair<ReduceGroup, uint32_t>& e) { return std::min<uint32_t>(e.first.count, 1u << 16); });
  for (size_t i = 0; i < g.size(); i++) {
    g[i] = z[i].first;
    out[i] = z[i].second;
  }
}

// SHA-256 compressions of expand_message_xmd(msg, DST, 48) for a message of
// `len` bytes: b_0 over Z_pad (64) || msg || I2OSP(48, 2) || 0x00 || DST',
// then b_1 and b_2 over 32 + 1 + |DST'| bytes (RFC 9380 section 5.3.1).
static uint64_t xmd48_blocks(uint64_t len, uint64_t dst_len) {
  auto blocks = [](uint64_t n) { return (n + 9 + 63) / 64; };
  return blocks(64 + len + 3 + dst_len + 1) + 2 * blocks(32 + 1 + dst_len + 1);
}

// Work units of the plan (cubbs_hip_batch_work).
static void count_work(Batch& b, const Plan& p) {
  CubbsHipBatchWork& w = b.work;
  w.items = b.nv;
  w.comb_terms = b.comb_terms;
  w.comb_madds = b.comb_madds;
  w.glv_products = p.var.size();
  w.agg_products = p.agg_var.size();
  // multi-term products: the ungated jobs, then the gated ones (counted as
  // if every group failed until results say otherwise)
  uint64_t* u = b.ungated_straus;
  u[0] = u[1] = u[2] = u[3] = 0;
  for (const auto& v : p.var_sum) {
    u[0]++;
    u[1] += v.count;
    if (v.flags & VAR_SUM_DUAL) {
      u[2]++;
      u[3] += v.count;
    }
  }
  w.straus_sums = u[0];
  w.straus_terms = u[1];
  w.straus_short_sums = u[2];
  w.straus_short_terms = u[3];
  for (const auto& g : b.gated_work) {
    w.straus_sums += g.sums;
    w.straus_terms += g.sum_terms;
    w.straus_short_sums += g.duals;
    w.straus_short_terms += g.dual_terms;
  }
  w.decoded_points = p.dec_g1.size();
  uint64_t adds = 0;
  for (const auto& g : p.groups1) adds += g.count ? g.count - 1 : 0;
  for (const auto& g : p.groups2) adds += g.count ? g.count - 1 : 0;
  for (const auto& lv : p.pst_groups)
    for (const auto& g : lv) adds += g.count ? g.count - 1 : 0;
  b.ungated_point_adds = adds;
  for (const auto& g : b.gated_work) adds += g.adds;
  w.point_adds = adds;
  w.group_checks = p.n_group_pairs;
  w.pair_checks = p.pairs.size() - p.n_group_pairs;
  w.pair_checks_run = p.pairs.size();
  uint64_t sha = 0;
  // message hashes: their b_0 blocks after Z_pad, and b_1, b_2
  for (const auto& j : p.msg) sha += j.blocks + 2ull * p.tail_blocks;
  // domain input: PK || I2OSP(L, 8) || Q1 || H_1..H_L || api_id || I2OSP(|hdr|, 8) || hdr
  for (const auto& j : p.dom) sha += xmd48_blocks(96 + 8 + 48ull * (j.L + 1) + 44 + 8 + j.hdr_len, 48);
  // challenge: I2OSP(R, 8) || (I2OSP(i, 8) || m_i)_R || Abar, Bbar, D, T1, T2 || domain || I2OSP(|ph|, 8) || ph
  for (const auto& j : p.pvch) sha += xmd48_blocks(8 + 40ull * j.R + 240 + 32 + 8 + j.ph_len, 48);
  w.sha_blocks = sha;
}

void materialize(Batch& b, Plan& p) {
  // The lanes of a wave wait for the longest of their jobs: order message
  // hashes by SHA-256 block count and domains by input length (each job
  // names its output slot, so the order does not change any result).
  sort_jobs(p.msg, [](const MsgHashJob& j) { return std::min<uint32_t>(j.blocks, 1u << 16); });
  if (!p.msg.empty() && !p.tail_blocks) p.message_tails();
  sort_jobs(p.dom, [](const DomainJob& j) {
    return std::min<uint32_t>((48u * j.L + j.hdr_len) >> 6, 1u << 16);
  });
  // the same for comb chunks (a row's last chunk is usually short)
  sort_jobs(p.chunks, [](const CombChunk& c) { return std::min<uint32_t>(c.term_count, 1u << 16); });
  if (!p.fb_chunks.empty()) {
    std::vector<std::pair<CombChunk, uint32_t>> cg(p.fb_chunks.size());
    for (size_t i = 0; i < cg.size(); i++) cg[i] = {p.fb_chunks[i], p.fb_gate[i]};
    sort_jobs(cg, [](const std::pair<CombChunk, uint32_t>& c) {
      return std::min<uint32_t>(c.first.term_count, 1u << 16);
    });
    for (size_t i = 0; i < cg.size(); i++) {
      p.fb_chunks[i] = cg[i].first;
      p.fb_gate[i] = cg[i].second;
    }
  }
  sort_groups(p.groups1, p.out1);
  std::vector<uint8_t> H;
  H.reserve(p.blob.size() + 4096);
  auto place = [&](const void* src, size_t bytes) {
    size_t off = align256(H.size());
    H.resize(off + bytes);
    if (bytes) std::memcpy(H.data() + off, src, bytes);
    return off;
  };
  size_t o_msg = place(p.msg.data(), p.msg.size() * sizeof(MsgHashJob));
  size_t o_dom = place(p.dom.data(), p.dom.size() * sizeof(DomainJob));
  size_t o_dg1 = place(p.dec_g1.data(), p.dec_g1.size() * sizeof(DecodeG1Job));
  size_t o_dfr = place(p.dec_fr.data(), p.dec_fr.size() * sizeof(DecodeFrJob));
  size_t o_fill = place(p.fills.data(), p.fills.size() * sizeof(TermFill));
  // one product launch: the plan's products, then the aggregation products
  std::vector<VarJob> all_var(p.var);
  all_var.insert(all_var.end(), p.agg_var.begin(), p.agg_var.end());
  size_t o_var = place(all_var.data(), all_var.size() * sizeof(VarJob));
  // multi-term products: the longest first (ordered by number of terms and
  // dual or not, so that a wave's lanes run the same chains), each with its
  // own workspace; then the gated jobs
  auto sum_order = [](const VarSum& x, const VarSum& y) {
    const uint32_t kx = 2 * x.count + (x.flags & VAR_SUM_DUAL), ky = 2 * y.count + (y.flags & VAR_SUM_DUAL);
    return kx > ky;
  };
  std::stable_sort(p.var_sum.begin(), p.var_sum.end(), sum_order);
  std::stable_sort(p.fb_sum.begin(), p.fb_sum.end(), sum_order);
  uint64_t sum_ws_words = 0;
  for (auto& v : p.var_sum) {
    v.ws = (uint32_t)sum_ws_words;
    sum_ws_words += var_sum_ws_words(v.count);
  }
  for (auto& v : p.fb_sum) {
    v.ws = (uint32_t)sum_ws_words;
    sum_ws_words += var_sum_ws_words(v.count);
  }
  if (sum_ws_words >> 32) throw std::length_error("product workspace beyond 2^32 words");
  size_t o_vsum = place(p.var_sum.data(), p.var_sum.size() * sizeof(VarSum));
  size_t o_fbs = place(p.fb_sum.data(), p.fb_sum.size() * sizeof(VarSum));
  std::vector<uint32_t> fb_sum_gate(p.fb_sum.size()), fb_red_gate(p.fb_red.size());
  for (size_t i = 0; i < p.fb_sum.size(); i++) fb_sum_gate[i] = p.fb_sum[i].gate;
  for (size_t i = 0; i < p.fb_red.size(); i++) fb_red_gate[i] = p.fb_red[i].gate;
  size_t o_fbsg = place(fb_sum_gate.data(), fb_sum_gate.size() * 4);
  size_t o_fbrg = place(fb_red_gate.data(), fb_red_gate.size() * 4);
  size_t o_chk = place(p.chunks.data(), p.chunks.size() * sizeof(CombChunk));
  size_t o_tg = place(p.term_gen.data(), p.term_gen.size() * 4);
  size_t o_g1 = place(p.groups1.data(), p.groups1.size() * sizeof(ReduceGroup));
  size_t o_o1 = place(p.out1.data(), p.out1.size() * 4);
  size_t o_g2 = place(p.groups2.data(), p.groups2.size() * sizeof(ReduceGroup));
  size_t o_o2 = place(p.out2.data(), p.out2.size() * 4);
  size_t o_pair = place(p.pairs.data(), p.pairs.size() * sizeof(PairJob));
  size_t o_sp = place(p.sum_pairs.data(), p.sum_pairs.size() * 4);
  size_t o_fbc = place(p.fb_chunks.data(), p.fb_chunks.size() * sizeof(CombChunk));
  size_t o_fbg = place(p.fb_gate.data(), p.fb_gate.size() * 4);
  size_t o_fbr = place(p.fb_red.data(), p.fb_red.size() * sizeof(GatedReduce));
  std::vector<std::pair<size_t, size_t>> o_lv;
  for (size_t l = 0; l < p.pst_groups.size(); l++)
    o_lv.emplace_back(place(p.pst_groups[l].data(), p.pst_groups[l].size() * sizeof(ReduceGroup)),
                      place(p.pst_out[l].data(), p.pst_out[l].size() * 4));
  size_t o_comp = place(p.comp.data(), p.comp.size() * sizeof(CompressJob));
  size_t o_pipt = place(p.pip_terms.data(), p.pip_terms.size() * sizeof(PipTerm));
  size_t o_pipo = place(p.pip_out.data(), p.pip_out.size() * 4);
  size_t o_sign = place(p.sign.data(), p.sign.size() * sizeof(SignJob));
  size_t o_pvch = place(p.pvch.data(), p.pvch.size() * sizeof(PvChallengeJob));
  size_t o_pg = place(p.pg.data(), p.pg.size() * sizeof(PgJob));
  size_t o_di = place(p.disc_index.data(), p.disc_index.size() * 4);
  size_t o_ds = place(p.disc_scalar.data(), p.disc_scalar.size() * 4);
  size_t o_us = place(p.und_scalar.data(), p.und_scalar.size() * 4);
  size_t o_dst = place(p.map_dst.data(), p.map_dst.size());
  size_t o_blob = place(p.blob.data(), std::max<size_t>(p.blob.size(), 1));
  b.in_bytes = align256(H.size());
  Arena* ar = tl_arena;
  b.from_arena = ar != nullptr;
  if (ar) {
    b.h_in = ar->alloc_host(b.in_bytes);
    b.d_in = ar->alloc_dev(b.in_bytes);
  } else {
    check(hipHostMalloc(&b.h_in, b.in_bytes, hipHostMallocDefault));
    check(hipMalloc(&b.d_in, b.in_bytes));
  }
  std::memcpy(b.h_in, H.data(), H.size());
  uint8_t* D = b.d_in;
  b.msg = reinterpret_cast<MsgHashJob*>(D + o_msg);
  b.n_msg = (uint32_t)p.msg.size();
  b.dom = reinterpret_cast<DomainJob*>(D + o_dom);
  b.n_dom = (uint32_t)p.dom.size();
  b.dec_g1 = reinterpret_cast<DecodeG1Job*>(D + o_dg1);
  b.n_dec_g1 = (uint32_t)p.dec_g1.size();
  b.dec_fr = reinterpret_cast<DecodeFrJob*>(D + o_dfr);
  b.n_dec_fr = (uint32_t)p.dec_fr.size();
  b.fills = reinterpret_cast<TermFill*>(D + o_fill);
  b.n_fills = (uint32_t)p.fills.size();
  b.var = reinterpret_cast<VarJob*>(D + o_var);
  b.n_var = (uint32_t)all_var.size();
  b.var_sum = reinterpret_cast<VarSum*>(D + o_vsum);
  b.n_var_sum = (uint32_t)p.var_sum.size();
  b.chunks = reinterpret_cast<CombChunk*>(D + o_chk);
  b.n_chunks = (uint32_t)p.chunks.size();
  b.term_gen = reinterpret_cast<uint32_t*>(D + o_tg);
  b.n_terms = (uint32_t)p.term_gen.size();
  b.groups1 = reinterpret_cast<ReduceGroup*>(D + o_g1);
  b.out1 = reinterpret_cast<uint32_t*>(D + o_o1);
  b.n_groups1 = (uint32_t)p.groups1.size();
  b.groups2 = reinterpret_cast<ReduceGroup*>(D + o_g2);
  b.out2 = reinterpret_cast<uint32_t*>(D + o_o2);
  b.n_groups2 = (uint32_t)p.groups2.size();
  b.pairs = reinterpret_cast<PairJob*>(D + o_pair);
  b.n_pairs = (uint32_t)p.pairs.size();
  b.sum_pairs = reinterpret_cast<uint32_t*>(D + o_sp);
  b.fb_chunks = reinterpret_cast<CombChunk*>(D + o_fbc);
  b.fb_gate = reinterpret_cast<uint32_t*>(D + o_fbg);
  b.n_fb_chunks = (uint32_t)p.fb_chunks.size();
  b.fb_red = reinterpret_cast<GatedReduce*>(D + o_fbr);
  b.n_fb_red = (uint32_t)p.fb_red.size();
  b.fb_sum = reinterpret_cast<VarSum*>(D + o_fbs);
  b.n_fb_sum = (uint32_t)p.fb_sum.size();
  b.fb_sum_gate = reinterpret_cast<uint32_t*>(D + o_fbsg);
  b.fb_red_gate = reinterpret_cast<uint32_t*>(D + o_fbrg);
  b.n_agg_slots = p.n_agg_slots;
  b.n_spairs = p.n_super_pairs;
  b.n_gpairs = p.n_group_pairs - p.n_super_pairs;
  for (size_t l = 0; l < o_lv.size(); l++)
    b.levels.push_back(Batch::Level{reinterpret_cast<ReduceGroup*>(D + o_lv[l].first),
                                    reinterpret_cast<uint32_t*>(D + o_lv[l].second),
                                    (uint32_t)p.pst_groups[l].size()});
  b.pip_terms = reinterpret_cast<PipTerm*>(D + o_pipt);
  b.pip_out = reinterpret_cast<uint32_t*>(D + o_pipo);
  if (!p.pip_terms.empty()) b.pip = pip_shape((uint32_t)p.pip_terms.size(), (uint32_t)p.pip_out.size());
  b.comp = reinterpret_cast<CompressJob*>(D + o_comp);
  b.n_comp = (uint32_t)p.comp.size();
  b.sign = reinterpret_cast<SignJob*>(D + o_sign);
  b.n_sign = (uint32_t)p.sign.size();
  b.pvch = reinterpret_cast<PvChallengeJob*>(D + o_pvch);
  b.n_pvch = (uint32_t)p.pvch.size();
  b.pg = reinterpret_cast<PgJob*>(D + o_pg);
  b.n_pg = (uint32_t)p.pg.size();
  b.disc_index = reinterpret_cast<uint32_t*>(D + o_di);
  b.disc_scalar = reinterpret_cast<uint32_t*>(D + o_ds);
  b.und_scalar = reinterpret_cast<uint32_t*>(D + o_us);
  b.map_dst = D + o_dst;
  b.map_dst_len = (uint32_t)p.map_dst.size();
  b.blob = D + o_blob;
  b.tail = reinterpret_cast<const uint32_t*>(b.blob + p.tail_off);
  b.tail_blocks = p.tail_blocks;
  b.seeded = p.seeded;
  b.seeds = D + o_blob + p.seeds_off;
  auto chunk_madds = [&](const CombChunk& c) {
    uint32_t m = 0;
    for (uint32_t t = c.term_begin; t < c.term_begin + c.term_count; t++)
      m += p.term_gen[t] < b.n16 ? COMB16_WINDOWS : COMB_WINDOWS;
    return m;
  };
  for (const auto& c : p.chunks) {
    b.comb_terms += c.term_count;
    b.comb_madds += chunk_madds(c);
  }
  b.ungated_comb_terms = b.comb_terms;
  b.ungated_comb_madds = b.comb_madds;
  // gated work, counted as if every group failed until results say otherwise
  for (size_t i = 0; i < p.fb_chunks.size(); i++) {
    const uint32_t m = chunk_madds(p.fb_chunks[i]);
    b.gated_work.push_back(Batch::GatedWork{p.fb_gate[i], p.fb_chunks[i].term_count, m, 0, 0, 0, 0, 0});
    b.comb_terms += p.fb_chunks[i].term_count;
    b.comb_madds += m;
  }
  for (const auto& r : p.fb_red) {
    const uint32_t in = r.count + (r.extra != REDUCE_NO_EXTRA ? 1u : 0u);
    b.gated_work.push_back(Batch::GatedWork{r.gate, 0, 0, in ? in - 1 : 0, 0, 0, 0, 0});
  }
  for (const auto& v : p.fb_sum) {
    const bool dual = (v.flags & VAR_SUM_DUAL) != 0;
    b.gated_work.push_back(Batch::GatedWork{v.gate, 0, 0, 0, 1, v.count, dual ? 1u : 0u, dual ? v.count : 0u});
  }
  count_work(b, p);
  if (p.n_agg_slots) {
    b.pair_group.resize(p.pairs.size() - p.n_super_pairs);
    for (size_t i = p.n_super_pairs; i < p.pairs.size(); i++)
      b.pair_group[i - p.n_super_pairs] =
          (p.pairs[i].flags & PAIR_GROUPED) ? p.pairs[i].group : PRIM_NO_SLOT;
  }
  // work area
  size_t w = 0;
  auto take = [&](size_t bytes) {
    size_t off = w;
    w = align256(w + bytes);
    return off;
  };
  size_t o_sc = take(4ull * FR_WORDS * std::max<uint32_t>(p.n_scalars, 1));
  size_t o_pt = take(4ull * G1A_WORDS * std::max<uint32_t>(p.n_points, 1));
  size_t o_pi = take(std::max<uint32_t>(p.n_points, 1));
  size_t o_pa = take(4ull * G1J_WORDS * std::max<uint32_t>(p.n_partials, 1));
  const uint32_t slots = b.nv + p.n_agg_slots;  // items, then aggregated groups
  size_t o_st = take(std::max<uint32_t>(slots, 1));
  size_t o_ga = take(std::max<uint32_t>(slots, 1));
  size_t o_va = take(std::max<uint32_t>(slots, 1));
  size_t o_out = take(std::max<size_t>(p.out_bytes, 1));
  size_t o_l2 = take(4ull * PAIR_L2_WORDS * std::max<size_t>(p.pairs.size(), 1));
  size_t o_plist = take(4ull * std::max<size_t>(p.pairs.size(), 1));
  size_t o_pairn = take(4);
  // product workspaces: the plan's jobs, then the aggregation jobs
  size_t o_vt = take(4ull * VAR_WS_WORDS * std::max<size_t>(all_var.size(), 1));
  size_t o_glist = take(4ull * std::max<size_t>(p.fb_sum.size() + p.fb_chunks.size() + p.fb_red.size(), 1));
  size_t o_gcnt = take(16);
  size_t o_sws = take(4ull * std::max<uint64_t>(sum_ws_words, 1));
  const PipShape& ps = b.pip;
  const size_t pip_on = ps.terms ? 1 : 0;
  size_t o_pcnt = take(4ull * (ps.nbt + 1) * pip_on);
  size_t o_poff = take(4ull * (ps.nbt + 1) * pip_on);
  size_t o_pcur = take(4ull * ps.nbt * pip_on);
  size_t o_ppof = take(4ull * (ps.nbt + 1) * pip_on);
  size_t o_pmap = take(4ull * ps.max_pieces * pip_on);
  size_t o_psrt = take(4ull * ps.max_entries * pip_on);
  size_t o_phdr = take(16 * pip_on);
  size_t o_ppcs = take(4ull * G1J_WORDS * ps.max_pieces * pip_on);
  size_t o_pbkt = take(4ull * G1J_WORDS * ps.nbt * pip_on);
  b.work_bytes = w;
  if (ar)
    b.d_work = ar->alloc_dev(w);
  else
    check(hipMalloc(&b.d_work, w));
  uint8_t* W = b.d_work;
  b.scalars = reinterpret_cast<uint32_t*>(W + o_sc);
  b.points = reinterpret_cast<uint32_t*>(W + o_pt);
  b.pinf = W + o_pi;
  b.partials = reinterpret_cast<uint32_t*>(W + o_pa);
  b.status = W + o_st;
  b.gate = W + o_ga;
  b.valid = W + o_va;
  b.out = W + o_out;
  b.l2buf = reinterpret_cast<uint32_t*>(W + o_l2);
  b.pair_list = reinterpret_cast<uint32_t*>(W + o_plist);
  b.pair_count = reinterpret_cast<uint32_t*>(W + o_pairn);
  b.vtab = reinterpret_cast<uint32_t*>(W + o_vt);
  b.sum_ws = reinterpret_cast<uint32_t*>(W + o_sws);
  b.gate_list = reinterpret_cast<uint32_t*>(W + o_glist);
  b.gate_count = reinterpret_cast<uint32_t*>(W + o_gcnt);
  if (pip_on) {
    auto u32 = [&](size_t o) { return reinterpret_cast<uint32_t*>(W + o); };
    b.pipb = PipBuffers{u32(o_pcnt), u32(o_poff), u32(o_pcur), u32(o_ppof), u32(o_pmap),
                        u32(o_psrt), u32(o_phdr), u32(o_ppcs), u32(o_pbkt)};
  }
  b.out_bytes = p.out_bytes;
  b.n_scalars = p.n_scalars;
  b.n_points = p.n_points;
  b.n_partials = p.n_partials;
  if (ar) {
    b.stream = ar->streams[ar->cur];
    b.aux = ar->aux[ar->cur];
    b.ev_fork = ar->fork[ar->cur];
    b.ev_join = ar->join[ar->cur];
  } else {
    {
      Context& c = ctx();
      std::lock_guard<std::mutex> g(c.mu);
      b.stream = next_stream_locked(c);
      b.aux = tl_single_stream ? nullptr : next_stream_locked(c);
    }
    check(hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming));
    check(hipEventCreateWithFlags(&b.ev_join, hipEventDisableTiming));
  }
  b.res_stride = align256(std::max<uint32_t>(slots, 1));
  const size_t res_bytes = 2 * b.res_stride + align256(std::max<size_t>(p.out_bytes, 1));
  if (ar)
    b.h_res = ar->alloc_host(res_bytes);
  else
    check(hipHostMalloc(&b.h_res, res_bytes, hipHostMallocDefault));
  launch_copy(b.stream, b.d_in, b.h_in, b.in_bytes);
}

// ------------------------------------------------------------------ aggregated pairing checks
// Random linear combination of same-key checks (the batch test of the
// reference's verify_homogeneous_pre_b_aggregate, crates/cubbs-bbs/src/
// execution.rs:548-838, applied per group of one key): for checks
// e(a_i, PK) e(b_i, -G2) = 1, i in a group, the combined check
// e(sum c_i a_i, PK) e(sum c_i b_i, -G2) = 1 runs first, with coefficients
// c_i = s_i + t_i lambda (s_i, t_i uniform 64-bit from the OS CSPRNG, drawn
// after the inputs are fixed: 2^128 distinct values, and the GLV split of c_i
// is (s_i, t_i), so each product is a 64-bit ladder).  If it holds, every
// item's equation holds except with probability 2^-128 and the item is
// accepted without its own pairing; if not, each item is checked on its own.
// The per-item results are those of the individual checks either way.
// Off unless the preparing call asks for it (the prepare flags): a pairing
// launch costs about one check's latency however few checks it holds, so the
// extra combined-check launch (and the re-check launch when a group fails)
// outweighs the pairings saved unless several batches are in flight.
static thread_local uint32_t tl_agg_group = 0, tl_agg_super = 0, tl_agg_straus = 0;
void set_thread_agg_group(uint32_t n) { tl_agg_group = n; }
void set_thread_agg_super(uint32_t n) { tl_agg_super = n; }
void set_thread_agg_straus(uint32_t n) { tl_agg_straus = n; }
uint32_t agg_group_max() { return tl_agg_group; }
uint32_t agg_super_max() { return tl_agg_super; }
uint32_t agg_straus() { return tl_agg_straus; }

static void agg_coefficient(uint8_t out[32], uint64_t s, uint64_t t) {
  // s + t * lambda, lambda = GLV_LAMBDA (127.4 bits): below 2^192 < r
  const uint64_t l0 = (uint64_t)bls::GLV_LAMBDA[0] | (uint64_t)bls::GLV_LAMBDA[1] << 32;
  const uint64_t l1 = (uint64_t)bls::GLV_LAMBDA[2] | (uint64_t)bls::GLV_LAMBDA[3] << 32;
  const unsigned __int128 p0 = (unsigned __int128)t * l0 + s;
  const unsigned __int128 p1 = (unsigned __int128)t * l1 + (uint64_t)(p0 >> 64);
  const uint64_t w[4] = {(uint64_t)p0, (uint64_t)p1, (uint64_t)(p1 >> 64), 0};
  for (int k = 0; k < 4; k++)
    for (int j = 0; j < 8; j++) out[31 - 8 * k - j] = (uint8_t)(w[k] >> (8 * j));
}

void aggregate_pairs(Plan& p, uint32_t nv, uint32_t group_max, uint32_t super_max) {
  if (group_max < 2 || group_max > 8) return;
  if (super_max < 2 || super_max > 8) super_max = 1;
  std::vector<PairJob> in = std::move(p.pairs);
  p.pairs.clear();
  std::vector<uint32_t> order(in.size());
  for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
  auto plain = [&](const PairJob& j) { return j.pk_b == PAIR_NEG_G2 && j.flags == 0; };
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    const bool px = plain(in[x]), py = plain(in[y]);
    if (px != py) return px;
    return px && in[x].pk_slot < in[y].pk_slot;
  });
  std::vector<uint8_t> rnd(16 * in.size());
  fill_os_random(rnd.data(), rnd.size());
  std::vector<PairJob> supers, groups, grouped, single;
  std::vector<ReduceGroup> super_red;
  std::vector<uint32_t> super_out;
  // the groups of one key: runs of at most group_max checks (a key's lone
  // item stays a single check; its last lone item joins the previous run,
  // which then holds group_max + 1), taken super_max at a time
  std::vector<std::pair<size_t, size_t>> run;  // [begin, end) in order[]
  auto flush = [&]() {
    for (size_t c0 = 0; c0 < run.size(); c0 += super_max) {
      const size_t c1 = std::min(run.size(), c0 + super_max);
      const bool sup = c1 - c0 >= 2;
      const uint32_t total = (uint32_t)(run[c1 - 1].second - run[c0].first);
      const size_t first = run[c0].first;
      // the partials of the chunk's a points, then of its b points, contiguous
      uint32_t fa, fb;
      {
        std::vector<uint32_t> sc(total);
        for (uint32_t k = 0; k < total; k++) {
          uint64_t s0, t0;
          std::memcpy(&s0, &rnd[16 * order[first + k]], 8);
          std::memcpy(&t0, &rnd[16 * order[first + k] + 8], 8);
          uint8_t c[32];
          agg_coefficient(c, s0 | 1, t0);
          sc[k] = p.scalars(1);
          p.dec_fr.push_back(DecodeFrJob{p.bytes(c, 32), sc[k], nv + p.n_agg_slots, 0u, 0u});
        }
        fa = p.n_partials;
        for (uint32_t k = 0; k < total; k++)
          p.agg_var.push_back(VarJob{in[order[first + k]].a, sc[k], p.partials(VAR_PARTIALS)});
        fb = p.n_partials;
        for (uint32_t k = 0; k < total; k++)
          p.agg_var.push_back(VarJob{in[order[first + k]].b, sc[k], p.partials(VAR_PARTIALS)});
      }
      const uint32_t pk = in[order[first]].pk_slot;
      uint32_t sslot = 0;
      if (sup) {
        sslot = nv + p.n_agg_slots++;
        const uint32_t la = p.points(1), lb = p.points(1);
        super_red.push_back(ReduceGroup{fa, VAR_PARTIALS * total});
        super_out.push_back(la);
        super_red.push_back(ReduceGroup{fb, VAR_PARTIALS * total});
        super_out.push_back(lb);
        supers.push_back(PairJob{sslot, la, lb, pk, PAIR_NEG_G2, 0, 0});
      }
      for (size_t r = c0; r < c1; r++) {
        const uint32_t gslot = nv + p.n_agg_slots++;
        const uint32_t off = (uint32_t)(run[r].first - first);
        const uint32_t cnt = (uint32_t)(run[r].second - run[r].first);
        const uint32_t la = p.points(1), lb = p.points(1);
        p.groups2.push_back(ReduceGroup{fa + VAR_PARTIALS * off, VAR_PARTIALS * cnt});
        p.out2.push_back(la);
        p.groups2.push_back(ReduceGroup{fb + VAR_PARTIALS * off, VAR_PARTIALS * cnt});
        p.out2.push_back(lb);
        groups.push_back(PairJob{gslot, la, lb, pk, PAIR_NEG_G2, sup ? PAIR_GROUPED : 0u, sslot});
        for (size_t k = run[r].first; k < run[r].second; k++) {
          PairJob g = in[order[k]];
          g.flags |= PAIR_GROUPED;
          g.group = gslot;
          grouped.push_back(g);
        }
      }
    }
    run.clear();
  };
  size_t i = 0;
  while (i < order.size()) {
    const PairJob& j0 = in[order[i]];
    size_t e = i + 1;
    if (plain(j0))
      while (e < order.size() && e - i < group_max && plain(in[order[e]]) &&
             in[order[e]].pk_slot == j0.pk_slot)
        e++;
    if (!run.empty() && (!plain(j0) || in[order[run[0].first]].pk_slot != j0.pk_slot)) flush();
    if (plain(j0) && e - i == 1 && !run.empty() && run.back().second == i) {
      run.back().second = e;  // a last lone item of the key joins the previous group
    } else if (!plain(j0) || e - i < 2) {
      for (size_t k = i; k < e; k++) single.push_back(in[order[k]]);
    } else {
      run.push_back({i, e});
    }
    i = e;
  }
  flush();
  // the supergroup sums after the group sums (the longer reductions share waves)
  p.groups2.insert(p.groups2.end(), super_red.begin(), super_red.end());
  p.out2.insert(p.out2.end(), super_out.begin(), super_out.end());
  p.n_super_pairs = (uint32_t)supers.size();
  p.n_group_pairs = (uint32_t)(supers.size() + groups.size());
  p.pairs = std::move(supers);
  p.pairs.insert(p.pairs.end(), groups.begin(), groups.end());
  p.pairs.insert(p.pairs.end(), grouped.begin(), grouped.end());
  p.pairs.insert(p.pairs.end(), single.begin(), single.end());
}

// Lock the context, make generators/public keys resident, snapshot pointers.
// When the budget of 8-bit tables or the key table cannot serve the batch,
// the batch gets private copies (freed with it).
void attach(Batch& b, uint32_t max_gen_slots, const std::vector<const uint8_t*>& pks,
                   std::vector<uint32_t>& pk_slot) {
  Context& c = ctx();
  std::lock_guard<std::mutex> g(c.mu);
  init_locked(c);
  const uint32_t gens = std::max<uint32_t>(max_gen_slots, 2);
  ensure_generators_locked(c, gens);
  hold_locked(c, b.hold);
  b.tables = c.gen_tables;
  if (gens > c.table_slots) {
    // slots [0, gens) privately: the resident tables copied, the rest built
    const size_t per = (size_t)COMB_TABLE_WORDS * 4;
    uint32_t* t = nullptr;
    check(hipMalloc(&t, (size_t)gens * per));
    b.owned.push_back(t);
    check(hipMemcpyAsync(t, c.gen_tables, (size_t)c.table_slots * per, hipMemcpyDeviceToDevice, c.setup));
    build_comb_tables(c, t, c.table_slots, gens - c.table_slots);
    b.tables = t;
  }
  // the key table is allocated by the first lookup: snapshot it afterwards
  const bool resident = lookup_pks_locked(c, pks, pk_slot, b.hold);
  b.pk_bytes = c.pk_bytes;
  b.pk_status = c.pk_status;
  b.pk_lines = c.pk_lines;
  if (!resident) {
    // more distinct keys than the table can hold at once: this batch's keys
    // are prepared privately
    std::unordered_map<std::string, uint32_t> idx;
    std::vector<const uint8_t*> keys;
    pk_slot.assign(pks.size(), 0);
    for (size_t i = 0; i < pks.size(); i++) {
      auto r = idx.emplace(std::string(reinterpret_cast<const char*>(pks[i]), 96), (uint32_t)keys.size());
      if (r.second) keys.push_back(pks[i]);
      pk_slot[i] = r.first->second;
    }
    const size_t n = keys.size();
    uint8_t *bytes = nullptr, *status = nullptr;
    uint32_t* lines = nullptr;
    check(hipMalloc(&bytes, n * 96));
    b.owned.push_back(bytes);
    check(hipMalloc(&status, n));
    b.owned.push_back(status);
    check(hipMalloc(&lines, n * PK_LINES_WORDS * 4));
    b.owned.push_back(lines);
    std::vector<uint32_t> slots(n);
    for (size_t i = 0; i < n; i++) slots[i] = (uint32_t)i;
    prepare_pks_locked(c, keys, slots, bytes, status, lines);
    b.pk_bytes = bytes;
    b.pk_status = status;
    b.pk_lines = lines;
  }
  b.tables16 = c.gen_tables16;
  b.n16 = c.wide_slots;
  b.gen_comp = c.gen_comp;
  b.neg_g2_lines = c.neg_g2_lines;
  b.sha_zpad = c.sha_zpad;
}

static std::string map_dst() { return std::string(kApiId) + "MAP_MSG_TO_SCALAR_AS_HASH_"; }

// ------------------------------------------------------------------ Verify
// Aggregated Verify.  Every item's check is scaled by its coefficient c
// (s + t lambda, s and t 64-bit from the OS CSPRNG, s odd so c != 0), as
// e(c A, PK) e(c (B - e A), -G2) = 1, which holds exactly when the item's own
// check does (c is invertible mod r).  The items of one key are grouped,
// ordered by message count, and a group of items i checks
//   e(sum c_i A_i, PK) e(sum c_i B_i - sum c_i e_i A_i, -G2) = 1.
// B_i = P1 + Q1 dom_i + sum_j H_j m_ij is linear in its scalars, so the
// group's sum c_i B_i is ONE fixed-base MSM over summed term scalars:
//   P1 sum c_i + Q1 sum c_i dom_i + sum_j H_j (sum over i with L_i >= j of c_i m_ij),
// max L_i + 2 terms for the group instead of sum (L_i + 2).  The A sides,
// sum c_i A_i and sum (-c_i e_i) A_i, are multi-term product jobs of up to
// VAR_TERMS_MAX items each: one table of multiples per point serves both
// sums, and each sum takes one chain of doublings for all its points (64
// for the coefficients' 64-bit GLV halves, 128 for the full scalars)
// instead of one chain per item and product.  An item's own products c A,
// (-c e) A and terms c, c dom, c m_j (filled as well) and the sums forming
// its check points run only when its group check fails (gated on the
// group's status slot, after the group checks).  A key's last item, when
// it is left alone, is checked on its own throughout.  Supergroups (runs of
// up to super_max groups of a key) are checked before their groups.
//
// Item k of `dev` is device item ibase + k; group status slots are numbered
// from slot_base (past every device item of the batch).
static void plan_verify_grouped(Plan& p, const CuBbsBbsGpuVerifyItem* items,
                                const std::vector<size_t>& dev, const uint32_t* pk_slot,
                                uint32_t ibase, uint32_t slot_base, uint32_t CH, uint32_t group_max,
                                uint32_t super_max, uint32_t straus) {
  const uint32_t nv = (uint32_t)dev.size();
  if (straus > VAR_TERMS_MAX) straus = VAR_TERMS_MAX;
  if (super_max < 2 || super_max > 8) super_max = 1;
  std::vector<uint32_t> Ls(nv);
  for (uint32_t k = 0; k < nv; k++) Ls[k] = (uint32_t)items[dev[k]].message_count;
  // groups: the items of one key ordered by message count, in runs of at
  // most group_max (a key's lone item is a single check; a last lone item
  // joins the previous run, which then holds group_max + 1)
  std::vector<uint32_t> order(nv);
  for (uint32_t k = 0; k < nv; k++) order[k] = k;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    if (pk_slot[x] != pk_slot[y]) return pk_slot[x] < pk_slot[y];
    return Ls[x] < Ls[y];
  });
  struct Run {
    uint32_t begin, end;  // into order[]
  };
  std::vector<std::vector<Run>> key_runs;  // per key, its groups
  std::vector<uint32_t> singles;
  for (uint32_t i = 0; i < nv;) {
    uint32_t e = i;
    while (e < nv && pk_slot[order[e]] == pk_slot[order[i]]) e++;
    key_runs.emplace_back();
    for (uint32_t r = i; r < e; r += group_max) {
      const uint32_t re = std::min(e, r + group_max);
      if (re - r >= 2)
        key_runs.back().push_back(Run{r, re});
      else if (!key_runs.back().empty())
        key_runs.back().back().end = re;  // a last lone item joins the previous group
      else
        singles.push_back(order[r]);
    }
    i = e;
  }
  // term slots first: [tbase, tbase + T) -- the groups' summed terms, then
  // every item's own terms
  uint32_t T = 0;
  for (const auto& runs : key_runs)
    for (const Run& r : runs) T += Ls[order[r.end - 1]] + 2;
  for (uint32_t k = 0; k < nv; k++) T += Ls[k] + 2;
  const uint32_t tbase = p.scalars(T);
  p.term_gen.resize(tbase + T);
  uint32_t term = tbase;
  // inputs of every item: A, c, r - c e, the domain and message scalars
  std::vector<uint8_t> rnd(16ull * nv);
  fill_os_random(rnd.data(), rnd.size());
  std::vector<uint32_t> pa(nv), sc(nv), se(nv), raw(nv);
  for (uint32_t k = 0; k < nv; k++) {
    const auto& it = items[dev[k]];
    const uint32_t L = Ls[k];
    const size_t sig = p.bytes(it.signature.data, 80);
    const size_t hdr = p.bytes(it.header.data, it.header.len);
    pa[k] = p.points(1);
    p.dec_g1.push_back(DecodeG1Job{sig, pa[k], ibase + k});
    uint64_t s0, t0;
    std::memcpy(&s0, &rnd[16ull * k], 8);
    std::memcpy(&t0, &rnd[16ull * k + 8], 8);
    uint8_t c[32];
    agg_coefficient(c, s0 | 1, t0);
    const size_t c_off = p.bytes(c, 32);
    sc[k] = p.scalars(1);
    se[k] = p.scalars(1);
    raw[k] = p.scalars(L + 1);
    p.dec_fr.push_back(DecodeFrJob{c_off, sc[k], ibase + k, 0u, 0u});
    p.dec_fr.push_back(DecodeFrJob{sig + 48, se[k], ibase + k, 1u | 4u, 0u, c_off});  // r - c e
    p.dom.push_back(DomainJob{hdr, (uint32_t)it.header.len, pk_slot[k], L, raw[k]});
    for (uint32_t m = 0; m < L; m++) p.hash_message(it.messages[m].data, it.messages[m].len, raw[k] + 1 + m);
  }
  // an item's own terms: c, c dom, c m_1 .. c m_L
  auto own_terms = [&](uint32_t k) {
    const uint32_t t0 = term, L = Ls[k];
    for (uint32_t t = 0; t < L + 2; t++) p.term_gen[t0 + t] = t;
    p.fills.push_back(TermFill{t0, sc[k], 0, 0});
    for (uint32_t t = 0; t <= L; t++) p.fills.push_back(TermFill{t0 + 1 + t, raw[k] + t, sc[k], 1});
    term += L + 2;
    return t0;
  };
  std::vector<PairJob> supers, groups, grouped, single;
  std::vector<std::pair<uint32_t, uint32_t>> xy;
  for (const auto& runs : key_runs) {
    for (size_t c0 = 0; c0 < runs.size(); c0 += super_max) {
      const size_t c1 = std::min(runs.size(), c0 + super_max);
      const bool sup = c1 - c0 >= 2;
      const uint32_t pk = pk_slot[order[runs[c0].begin]];
      uint32_t sslot = 0;
      if (sup) sslot = slot_base + p.n_agg_slots++;
      // the A sides: per item (straus == 0) the products c_i A_i and
      // (-c_i e_i) A_i, which the items' own checks reuse; else a group's
      // items in product jobs of at most `straus` points (nearly equal
      // parts), each forming the two sums of its points, with -c_i e_i and
      // with c_i (nj: a group's partials per side)
      std::vector<uint32_t> nj(c1 - c0);
      uint32_t jobs = 0;
      for (size_t r = c0; r < c1; r++) {
        const uint32_t cnt = runs[r].end - runs[r].begin;
        nj[r - c0] = straus ? (cnt + straus - 1) / straus : cnt;
        jobs += nj[r - c0];
      }
      // partials: the c_i A_i side of every group of the chunk, then per
      // group its (-c_i e_i) A_i side followed by its comb partials
      const uint32_t fx = p.partials(jobs);
      const uint32_t ybegin = p.n_partials;
      std::vector<uint32_t> xb(c1 - c0), vb(c1 - c0), vend(c1 - c0), gslot(c1 - c0);
      uint32_t xo = fx;
      for (size_t r = c0; r < c1; r++) {
        const Run& run = runs[r];
        const uint32_t cnt = run.end - run.begin, nr = nj[r - c0];
        xb[r - c0] = xo;
        vb[r - c0] = p.partials(nr);
        for (uint32_t q = 0; q < nr; q++) {
          if (!straus) {
            const uint32_t k = order[run.begin + q];
            p.agg_var.push_back(VarJob{pa[k], sc[k], xo + q});
            p.var.push_back(VarJob{pa[k], se[k], vb[r - c0] + q});
            continue;
          }
          const uint32_t a = run.begin + cnt * q / nr, e = run.begin + cnt * (q + 1) / nr;
          uint32_t jp[VAR_TERMS_MAX], js1[VAR_TERMS_MAX], js2[VAR_TERMS_MAX];
          for (uint32_t m = a; m < e; m++) {
            const uint32_t k = order[m];
            jp[m - a] = pa[k];
            js1[m - a] = se[k];
            js2[m - a] = sc[k];
          }
          p.var_dual(jp, js1, js2, e - a, vb[r - c0] + q, xo + q);
        }
        xo += nr;
        uint32_t lmax = 0;
        for (uint32_t m = run.begin; m < run.end; m++) lmax = std::max(lmax, Ls[order[m]]);
        const uint32_t gt = term;
        term += lmax + 2;
        for (uint32_t t = 0; t < lmax + 2; t++) {
          p.term_gen[gt + t] = t;
          xy.clear();
          for (uint32_t m = run.begin; m < run.end; m++) {
            const uint32_t k = order[m];
            if (t == 0)
              xy.push_back({sc[k], TERM_SUM_ONE});                 // P1: sum c_i
            else if (t - 1 <= Ls[k])
              xy.push_back({raw[k] + t - 1, sc[k]});               // Q1, H_j: sum c_i x_i
          }
          p.term_sum(gt + t, xy);
        }
        p.comb(gt, lmax + 2, CH);
        vend[r - c0] = p.n_partials;
        gslot[r - c0] = slot_base + p.n_agg_slots++;
      }
      const uint32_t yend = p.n_partials;
      // the sums, once the chunk's partials are all allocated (a long sum's
      // tree levels take partial slots after them)
      for (size_t r = c0; r < c1; r++) {
        const uint32_t la = p.points(1), lb = p.points(1);
        p.reduce_to(xb[r - c0], nj[r - c0], la);
        p.reduce_to(vb[r - c0], vend[r - c0] - vb[r - c0], lb);
        groups.push_back(PairJob{gslot[r - c0], la, lb, pk, PAIR_NEG_G2, sup ? PAIR_GROUPED : 0u, sslot});
      }
      if (sup) {
        const uint32_t la = p.points(1), lb = p.points(1);
        p.reduce_to(fx, jobs, la);
        p.reduce_to(ybegin, yend - ybegin, lb);
        supers.push_back(PairJob{sslot, la, lb, pk, PAIR_NEG_G2, 0, 0});
      }
      // the items' own checks, for when their group check fails: their
      // products (the per-item ones above, also with jobs of one item each,
      // else a gated dual product of c A and (-c e) A), and one gated sum
      // per check point over comb chunks of at least 4 terms (on this
      // critical path a longer chunk costs less than the serial additions of
      // the sum it saves)
      const uint32_t fch = std::max(CH, 4u);
      for (size_t r = c0; r < c1; r++) {
        const Run& run = runs[r];
        for (uint32_t m = run.begin; m < run.end; m++) {
          const uint32_t k = order[m], L = Ls[k];
          const uint32_t ft = own_terms(k);
          const uint32_t g = gslot[r - c0];
          uint32_t px = xb[r - c0] + (m - run.begin), py = vb[r - c0] + (m - run.begin);
          if (straus > 1) {
            px = p.partials(1);
            py = p.partials(1);
            p.var_dual(&pa[k], &se[k], &sc[k], 1, py, px, g);
          }
          const uint32_t cb = p.comb_gated(ft, L + 2, fch, g);
          const uint32_t pa2 = p.points(1), pb = p.points(1);
          p.fb_red.push_back(GatedReduce{px, 1, REDUCE_NO_EXTRA, g, pa2});          // c A
          p.fb_red.push_back(GatedReduce{cb, p.n_partials - cb, py, g, pb});        // c (B - e A)
          grouped.push_back(PairJob{ibase + k, pa2, pb, pk, PAIR_NEG_G2, PAIR_GROUPED, g});
        }
      }
    }
  }
  for (uint32_t k : singles) {
    const uint32_t L = Ls[k];
    const uint32_t st = own_terms(k);
    const uint32_t pair_a = p.points(1), pb = p.points(1);
    const uint32_t fx = p.partials(1);
    p.reduce_to(fx, 1, pair_a);                               // c A
    const uint32_t first = p.partials(1);
    if (straus) {
      p.var_dual(&pa[k], &se[k], &sc[k], 1, first, fx);       // (-c e) A, c A
    } else {
      p.agg_var.push_back(VarJob{pa[k], sc[k], fx});          // c A
      p.var.push_back(VarJob{pa[k], se[k], first});           // (-c e) A
    }
    p.comb(st, L + 2, CH);                                    // c B
    p.reduce_to(first, p.n_partials - first, pb);             // c (B - e A)
    single.push_back(PairJob{ibase + k, pair_a, pb, pk_slot[k], PAIR_NEG_G2, 0, 0});
  }
  if (term != tbase + T) throw std::logic_error("grouped verify plan: term slots do not add up");
  p.n_super_pairs = (uint32_t)supers.size();
  p.n_group_pairs = (uint32_t)(supers.size() + groups.size());
  p.pairs = std::move(supers);
  p.pairs.insert(p.pairs.end(), groups.begin(), groups.end());
  p.pairs.insert(p.pairs.end(), grouped.begin(), grouped.end());
  p.pairs.insert(p.pairs.end(), single.begin(), single.end());
}

// Host validation of Verify items: per-item arrays of the batch at [at, at + n),
// device items numbered from dbase; returns the device items (indices into
// items) and the largest message count.
static std::vector<size_t> intake_verify(Batch& b, const CuBbsBbsGpuVerifyItem* items, size_t n,
                                         size_t at, uint32_t dbase, uint32_t& maxL) {
  std::vector<size_t> dev;
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    bool ok = it.public_key.len == 96 && it.public_key.data && it.signature.len == 80 &&
              it.signature.data && slice_ok(it.header) &&
              (it.message_count == 0 || it.messages != nullptr) && it.message_count < (1u << 24);
    for (size_t m = 0; ok && m < it.message_count; m++) ok = slice_ok(it.messages[m]);
    if (!ok) {
      b.pre_code[at + i] = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      continue;
    }
    b.dev_index[at + i] = dbase + (uint32_t)dev.size();
    dev.push_back(i);
    maxL = std::max<uint32_t>(maxL, (uint32_t)it.message_count);
  }
  return dev;
}

// Per-item Verify checks e(A, PK) e(B - e A, -G2) = 1 (no aggregation); item
// k of `dev` is device item ibase + k.
static void plan_verify_plain(Plan& p, const CuBbsBbsGpuVerifyItem* items, const std::vector<size_t>& dev,
                              const uint32_t* pk_slot, uint32_t ibase, uint32_t CH) {
  const uint32_t nv = (uint32_t)dev.size();
  // term slots first: [tbase, tbase + T)
  uint32_t T = 0;
  for (uint32_t k = 0; k < nv; k++) T += (uint32_t)items[dev[k]].message_count + 2;
  const uint32_t tbase = p.scalars(T);
  p.term_gen.resize(tbase + T);
  uint32_t term = tbase;
  for (uint32_t k = 0; k < nv; k++) {
    const auto& it = items[dev[k]];
    const uint32_t L = (uint32_t)it.message_count;
    size_t sig = p.bytes(it.signature.data, 80);
    size_t hdr = p.bytes(it.header.data, it.header.len);
    uint32_t pa = p.points(1), pb = p.points(1), se = p.scalars(1);
    p.dec_g1.push_back(DecodeG1Job{sig, pa, ibase + k});
    for (uint32_t t = 0; t < L + 2; t++) p.term_gen[term + t] = t;
    p.dom.push_back(DomainJob{hdr, (uint32_t)it.header.len, pk_slot[k], L, term + 1});
    for (uint32_t m = 0; m < L; m++) {
      p.hash_message(it.messages[m].data, it.messages[m].len, term + 2 + m);
    }
    p.fills.push_back(TermFill{term, 0, 0, 2});                       // P1 scalar = 1
    p.dec_fr.push_back(DecodeFrJob{sig + 48, se, ibase + k, 1u, 0u});  // store r - e
    uint32_t first = p.n_partials;
    p.var.push_back(VarJob{pa, se, p.partials(VAR_PARTIALS)});  // (-e) A
    p.comb(term, L + 2, CH);                                  // B
    p.reduce_to(first, p.n_partials - first, pb);             // B - e A
    p.pairs.push_back(PairJob{ibase + k, pa, pb, pk_slot[k], PAIR_NEG_G2, 0});
    term += L + 2;
  }
}

// ------------------------------------------------------------------ ProofVerify
static bool indexes_strictly_increasing(const CuBbsBbsGpuIndexSlice& s, size_t bound) {
  if (s.len && !s.data) return false;
  for (size_t i = 0; i < s.len; i++) {
    if (s.data[i] >= bound) return false;
    if (i && s.data[i] <= s.data[i - 1]) return false;
  }
  return true;
}

// Host validation of ProofVerify items, as intake_verify.
static std::vector<size_t> intake_proof_verify(Batch& b, const CuBbsBbsGpuProofVerifyItem* items,
                                               size_t n, size_t at, uint32_t dbase, uint32_t& maxL) {
  std::vector<size_t> dev;
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    // protocol.rs:774-779: bad disclosed indexes or count mismatch -> Ok(false)
    if (!indexes_strictly_increasing(it.disclosed_indexes, it.total_message_count) ||
        it.disclosed_count != it.disclosed_indexes.len ||
        (it.disclosed_count && !it.disclosed_messages)) {
      b.pre_code[at + i] = CUBBS_BBS_GPU_SUCCESS;
      continue;
    }
    bool ok = it.public_key.len == 96 && it.public_key.data && slice_ok(it.header) &&
              slice_ok(it.presentation_header) && it.proof.data && it.proof.len >= 272 &&
              it.proof.len < (1ull << 30);
    for (size_t m = 0; ok && m < it.disclosed_count; m++) ok = slice_ok(it.disclosed_messages[m]);
    if (!ok) {
      b.pre_code[at + i] = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      continue;
    }
    const size_t U = (it.proof.len - 240) / 32 - 1;
    const size_t R = it.disclosed_count;
    const size_t L = U + R;
    // proof.rs:3027-3033: an index beyond L - 1 fails verification
    bool idx_ok = true;
    for (size_t r = 0; r < R; r++) idx_ok = idx_ok && it.disclosed_indexes.data[r] < L;
    if (!idx_ok) {
      b.pre_code[at + i] = CUBBS_BBS_GPU_SUCCESS;
      continue;
    }
    b.dev_index[at + i] = dbase + (uint32_t)dev.size();
    dev.push_back(i);
    maxL = std::max<uint32_t>(maxL, (uint32_t)L);
  }
  return dev;
}

// ProofVerify checks; item k of `dev` is device item ibase + k, group status
// slots (group >= 2: aggregated checks) numbered from slot_base.
static void plan_proof_verify_items(Plan& p, const CuBbsBbsGpuProofVerifyItem* items,
                                    const std::vector<size_t>& dev, const uint32_t* pk_slot,
                                    uint32_t ibase, uint32_t slot_base, uint32_t CH, uint32_t group,
                                    uint32_t super) {
  const uint32_t nv = (uint32_t)dev.size();
  uint32_t T = 0;
  for (uint32_t k = 0; k < nv; k++) {
    const auto& it = items[dev[k]];
    T += (uint32_t)((it.proof.len - 240) / 32 - 1 + it.disclosed_count) + 2;
  }
  const uint32_t tbase = p.scalars(T);
  p.term_gen.resize(tbase + T);
  uint32_t term = tbase;
  for (uint32_t k = 0; k < nv; k++) {
    const uint32_t item = ibase + k;
    const auto& it = items[dev[k]];
    const uint32_t U = (uint32_t)((it.proof.len - 240) / 32 - 1);
    const uint32_t R = (uint32_t)it.disclosed_count;
    const uint32_t L = U + R;
    size_t pr = p.bytes(it.proof.data, 240 + 32 * (size_t)(U + 1));
    size_t hdr = p.bytes(it.header.data, it.header.len);
    size_t ph = p.bytes(it.presentation_header.data, it.presentation_header.len);
    uint32_t pab = p.points(1), pbb = p.points(1), pd = p.points(1), pt1 = p.points(1),
             pt2 = p.points(1);
    p.dec_g1.push_back(DecodeG1Job{pr + 0, pab, item});
    p.dec_g1.push_back(DecodeG1Job{pr + 48, pbb, item});
    p.dec_g1.push_back(DecodeG1Job{pr + 96, pd, item});
    uint32_t sc = p.scalars(4 + U);  // e^, r1^, r3^, m^_1..m^_U, then c
    uint32_t s_ehat = sc, s_r1 = sc + 1, s_r3 = sc + 2, s_mhat = sc + 3, s_c = sc + 3 + U;
    p.dec_fr.push_back(DecodeFrJob{pr + 144, s_ehat, item, 0, 0});
    p.dec_fr.push_back(DecodeFrJob{pr + 176, s_r1, item, 0, 0});
    p.dec_fr.push_back(DecodeFrJob{pr + 208, s_r3, item, 0, 0});
    for (uint32_t u = 0; u <= U; u++)
      p.dec_fr.push_back(DecodeFrJob{pr + 240 + 32ull * u, s_mhat + u, item, 0, 0});
    uint32_t s_dom = p.scalars(1);
    p.dom.push_back(DomainJob{hdr, (uint32_t)it.header.len, pk_slot[k], L, s_dom});
    uint32_t s_m = p.scalars(R);
    uint32_t disc_begin = (uint32_t)p.disc_index.size();
    for (uint32_t r = 0; r < R; r++) {
      p.hash_message(it.disclosed_messages[r].data, it.disclosed_messages[r].len, s_m + r);
      p.disc_index.push_back((uint32_t)it.disclosed_indexes.data[r]);
      p.disc_scalar.push_back(s_m + r);
    }
    // T2 = c P1 + (c dom) Q1 + sum_R (c m_i) H_i + sum_U m^_j H_j + r3^ D   (proof.rs:3048-3066)
    for (uint32_t t = 0; t < L + 2; t++) p.term_gen[term + t] = t;
    p.fills.push_back(TermFill{term, s_c, 0, 0});
    p.fills.push_back(TermFill{term + 1, s_dom, s_c, 1});
    uint32_t r = 0, u = 0;
    for (uint32_t i = 0; i < L; i++) {
      if (r < R && it.disclosed_indexes.data[r] == i) {
        p.fills.push_back(TermFill{term + 2 + i, s_m + r, s_c, 1});
        r++;
      } else {
        p.fills.push_back(TermFill{term + 2 + i, s_mhat + u, 0, 0});
        u++;
      }
    }
    // T1 = c Bbar + e^ Abar + r1^ D as three single products: one
    // multi-term job (shared doublings) does 1/3 less work but takes twice
    // a product's latency, which a pipeline of a few batches in flight feels
    // more than the work (measured: 3.05 -> 3.22 ms per 20-step step)
    uint32_t f1 = p.n_partials;
    p.var.push_back(VarJob{pbb, s_c, p.partials(VAR_PARTIALS)});
    p.var.push_back(VarJob{pab, s_ehat, p.partials(VAR_PARTIALS)});
    p.var.push_back(VarJob{pd, s_r1, p.partials(VAR_PARTIALS)});
    p.reduce_to(f1, 3 * VAR_PARTIALS, pt1);
    uint32_t f2 = p.n_partials;
    p.var.push_back(VarJob{pd, s_r3, p.partials(VAR_PARTIALS)});
    p.comb(term, L + 2, CH);
    p.reduce_to(f2, p.n_partials - f2, pt2);
    p.pvch.push_back(PvChallengeJob{ph, (uint32_t)it.presentation_header.len, item, R, disc_begin,
                                    pab, pbb, pd, pt1, pt2, s_dom, s_c});
    p.pairs.push_back(PairJob{item, pab, pbb, pk_slot[k], PAIR_NEG_G2, 0});
    term += L + 2;
  }
  if (term != tbase + T) throw std::logic_error("proof verify plan: term slots do not add up");
  aggregate_pairs(p, slot_base, group, super);
}

// A Verify batch and a ProofVerify batch as one device batch: the Verify
// items are device items [0, nv), the ProofVerify items follow, and the
// group status slots of both follow all items.  Every stage launches once
// for the items of both.  Pairing checks keep the level order the launches
// need: the supergroup checks of both parts, then their group checks, then
// their per-item checks.
static std::unique_ptr<Batch> plan_verify_set(const CuBbsBbsGpuVerifyItem* vitems, size_t nv,
                                              const CuBbsBbsGpuProofVerifyItem* pitems, size_t np,
                                              int kind) {
  auto b = std::make_unique<Batch>();
  b->kind = kind;
  b->n_total = nv + np;
  b->pre_code.assign(nv + np, -1);
  b->pre_value.assign(nv + np, 0);
  b->dev_index.assign(nv + np, 0);
  uint32_t maxL = 0;
  const std::vector<size_t> dv = intake_verify(*b, vitems, nv, 0, 0, maxL);
  const std::vector<size_t> dp = intake_proof_verify(*b, pitems, np, nv, (uint32_t)dv.size(), maxL);
  b->nv = (uint32_t)(dv.size() + dp.size());
  if (!b->nv) return b;
  std::vector<const uint8_t*> pks;
  pks.reserve(b->nv);
  for (size_t i : dv) pks.push_back(vitems[i].public_key.data);
  for (size_t i : dp) pks.push_back(pitems[i].public_key.data);
  std::vector<uint32_t> pk_slot;
  attach(*b, maxL + 2, pks, pk_slot);
  Plan p;
  std::string d = map_dst();
  p.map_dst.assign(d.begin(), d.end());
  const uint32_t group = agg_group_max(), super = agg_super_max();
  const bool agg = group >= 2 && group <= 8;
  if (!dv.empty()) {
    const uint32_t CH = pipeline_comb_chunk(0, nv);
    if (agg)
      plan_verify_grouped(p, vitems, dv, pk_slot.data(), 0, b->nv, CH, group, super, agg_straus());
    else
      plan_verify_plain(p, vitems, dv, pk_slot.data(), 0, CH);
  }
  std::vector<PairJob> vpairs = std::move(p.pairs);
  p.pairs.clear();
  const uint32_t vs = p.n_super_pairs, vg = p.n_group_pairs;
  p.n_super_pairs = p.n_group_pairs = 0;
  if (!dp.empty())
    plan_proof_verify_items(p, pitems, dp, pk_slot.data() + dv.size(), (uint32_t)dv.size(), b->nv,
                            pipeline_comb_chunk(1, np), agg ? group : 0u, super);
  if (!vpairs.empty()) {
    std::vector<PairJob> pp = std::move(p.pairs);
    const uint32_t ps = p.n_super_pairs, pg = p.n_group_pairs;
    p.pairs.clear();
    auto take = [&](const std::vector<PairJob>& v, size_t from, size_t to) {
      p.pairs.insert(p.pairs.end(), v.begin() + from, v.begin() + to);
    };
    take(vpairs, 0, vs);
    take(pp, 0, ps);
    take(vpairs, vs, vg);
    take(pp, ps, pg);
    take(vpairs, vg, vpairs.size());
    take(pp, pg, pp.size());
    p.n_super_pairs = vs + ps;
    p.n_group_pairs = vg + pg;
  }
  materialize(*b, p);
  return b;
}

std::unique_ptr<Batch> plan_verify(const CuBbsBbsGpuVerifyItem* items, size_t n) {
  return plan_verify_set(items, n, nullptr, 0, CUBBS_HIP_BATCH_VERIFY);
}

std::unique_ptr<Batch> plan_proof_verify(const CuBbsBbsGpuProofVerifyItem* items, size_t n) {
  return plan_verify_set(nullptr, 0, items, n, CUBBS_HIP_BATCH_PROOF_VERIFY);
}

std::unique_ptr<Batch> plan_verify_and_proof_verify(const CuBbsBbsGpuVerifyItem* vitems, size_t nv,
                                                    const CuBbsBbsGpuProofVerifyItem* pitems,
                                                    size_t np) {
  return plan_verify_set(vitems, nv, pitems, np, KIND_VERIFY_SET);
}

// ------------------------------------------------------------------ Sign
std::unique_ptr<Batch> plan_sign(const CuBbsBbsGpuSignItem* items, size_t n) {
  auto b = std::make_unique<Batch>();
  b->kind = KIND_SIGN;
  b->n_total = n;
  b->pre_code.assign(n, -1);
  b->pre_value.assign(n, 0);
  b->dev_index.assign(n, 0);
  std::vector<size_t> dev;
  uint32_t maxL = 0;
  uint64_t est_terms = 0;  // fixed-base terms of the batch (tuning class)
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    bool ok = it.secret_key.len == 32 && it.secret_key.data && it.public_key.len == 96 &&
              it.public_key.data && slice_ok(it.header) &&
              (it.message_count == 0 || it.messages != nullptr) && it.message_count < (1u << 24);
    for (size_t m = 0; ok && m < it.message_count; m++) ok = slice_ok(it.messages[m]);
    if (!ok) {
      b->pre_code[i] = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      continue;
    }
    if (!it.signature_out.data || it.signature_out.capacity < 80) {
      b->pre_code[i] = CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT;
      continue;
    }
    b->dev_index[i] = (uint32_t)dev.size();
    dev.push_back(i);
    maxL = std::max<uint32_t>(maxL, (uint32_t)it.message_count);
    est_terms += it.message_count + 2;
  }
  b->nv = (uint32_t)dev.size();
  if (!b->nv) return b;
  std::vector<const uint8_t*> pks(b->nv);
  for (uint32_t k = 0; k < b->nv; k++) pks[k] = items[dev[k]].public_key.data;
  std::vector<uint32_t> pk_slot;
  attach(*b, maxL + 2, pks, pk_slot);
  Plan p;
  std::string d = map_dst();
  p.map_dst.assign(d.begin(), d.end());
  const uint32_t CH = comb_chunk_terms(est_terms);
  uint32_t T = 0;
  for (uint32_t k = 0; k < b->nv; k++) T += (uint32_t)items[dev[k]].message_count + 2;
  p.scalars(T);
  p.term_gen.resize(T);
  uint32_t term = 0;
  for (uint32_t k = 0; k < b->nv; k++) {
    const auto& it = items[dev[k]];
    const uint32_t L = (uint32_t)it.message_count;
    size_t sk = p.bytes(it.secret_key.data, 32);
    size_t hdr = p.bytes(it.header.data, it.header.len);
    p.dom.push_back(DomainJob{hdr, (uint32_t)it.header.len, pk_slot[k], L, term + 1});
    for (uint32_t m = 0; m < L; m++) {
      p.hash_message(it.messages[m].data, it.messages[m].len, term + 2 + m);
    }
    for (uint32_t t = 0; t < L + 2; t++) p.term_gen[term + t] = t;
    uint32_t se = p.scalars(1), pa = p.points(1);
    uint32_t first = p.n_partials;
    p.comb(term, L + 2, CH);
    p.reduce_to(first, p.n_partials - first, pa);
    p.sign.push_back(SignJob{sk, k, term, L, se, pa, pk_slot[k]});
    term += L + 2;
  }
  p.out_bytes = 80ull * b->nv;
  materialize(*b, p);
  return b;
}

// ------------------------------------------------------------------ ProofGen
// scalars_mode: 0 = OS randomness (seeds drawn here), 1 = seeds given,
// 2 = explicit scalars.
std::unique_ptr<Batch> plan_proof_gen(const CuBbsBbsGpuProofGenItem* items, size_t n, int mode,
                                      const uint8_t* seeds32, const uint8_t* scalars_be,
                                      const size_t* scalar_offsets) {
  auto b = std::make_unique<Batch>();
  b->kind = KIND_PROOF_GEN;
  b->n_total = n;
  b->pre_code.assign(n, -1);
  b->pre_value.assign(n, 0);
  b->dev_index.assign(n, 0);
  std::vector<size_t> dev;
  uint32_t maxL = 0;
  uint64_t est_terms = 0;  // fixed-base terms of the batch (tuning class)
  for (size_t i = 0; i < n; i++) {
    const auto& it = items[i];
    bool ok = it.public_key.len == 96 && it.public_key.data && it.signature.len == 80 &&
              it.signature.data && slice_ok(it.header) && slice_ok(it.presentation_header) &&
              (it.message_count == 0 || it.messages != nullptr) && it.message_count < (1u << 24) &&
              indexes_strictly_increasing(it.disclosed_indexes, it.message_count);
    for (size_t m = 0; ok && m < it.message_count; m++) ok = slice_ok(it.messages[m]);
    if (!ok) {
      b->pre_code[i] = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      continue;
    }
    size_t U = it.message_count - it.disclosed_indexes.len;
    if (!it.proof_out.data || it.proof_out.capacity < 272 + 32 * U) {
      b->pre_code[i] = CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT;
      continue;
    }
    b->dev_index[i] = (uint32_t)dev.size();
    dev.push_back(i);
    maxL = std::max<uint32_t>(maxL, (uint32_t)it.message_count);
    est_terms += it.message_count + 2;
  }
  b->nv = (uint32_t)dev.size();
  if (!b->nv) return b;
  std::vector<const uint8_t*> pks(b->nv);
  for (uint32_t k = 0; k < b->nv; k++) pks[k] = items[dev[k]].public_key.data;
  std::vector<uint32_t> pk_slot;
  attach(*b, maxL + 2, pks, pk_slot);
  Plan p;
  std::string d = map_dst();
  p.map_dst.assign(d.begin(), d.end());
  const uint32_t CH = comb_chunk_terms(est_terms);
  uint32_t T = 0;
  for (uint32_t k = 0; k < b->nv; k++) {
    const auto& it = items[dev[k]];
    T += (uint32_t)it.message_count + 2 + (uint32_t)(it.message_count - it.disclosed_indexes.len);
  }
  p.scalars(T);
  p.term_gen.resize(T);
  std::vector<uint8_t> seeds;
  if (mode != 2) {
    seeds.resize(32ull * b->nv);
    if (mode == 1) {
      for (uint32_t k = 0; k < b->nv; k++) std::memcpy(seeds.data() + 32ull * k, seeds32 + 32 * dev[k], 32);
    } else {
      fill_os_random(seeds.data(), seeds.size());
    }
  }
  uint32_t term = 0;
  size_t out = 0;
  for (uint32_t k = 0; k < b->nv; k++) {
    const auto& it = items[dev[k]];
    const uint32_t L = (uint32_t)it.message_count;
    const uint32_t R = (uint32_t)it.disclosed_indexes.len;
    const uint32_t U = L - R;
    size_t sig = p.bytes(it.signature.data, 80);
    size_t hdr = p.bytes(it.header.data, it.header.len);
    size_t ph = p.bytes(it.presentation_header.data, it.presentation_header.len);
    uint32_t pa = p.points(1), pb = p.points(1);
    uint32_t pd = p.points(1), pab = p.points(1), pbb = p.points(1), pt1 = p.points(1),
             pt2 = p.points(1);
    uint32_t s_e = p.scalars(1), s_dom = p.scalars(1), s_m = p.scalars(L), rs = p.scalars(5 + U),
             kk = p.scalars(7);
    p.dec_g1.push_back(DecodeG1Job{sig, pa, k});
    p.dec_fr.push_back(DecodeFrJob{sig + 48, s_e, k, 0, 0});
    p.dom.push_back(DomainJob{hdr, (uint32_t)it.header.len, pk_slot[k], L, s_dom});
    for (uint32_t m = 0; m < L; m++) {
      p.hash_message(it.messages[m].data, it.messages[m].len, s_m + m);
    }
    if (mode == 2) {
      size_t so = p.bytes(scalars_be + 32 * scalar_offsets[dev[k]], 32ull * (5 + U));
      for (uint32_t t = 0; t < 5 + U; t++) p.dec_fr.push_back(DecodeFrJob{so + 32ull * t, rs + t, k, 0, 0});
    }
    // B = P1 + Q1 dom + sum H_i m_i
    for (uint32_t t = 0; t < L + 2; t++) p.term_gen[term + t] = t;
    p.fills.push_back(TermFill{term, 0, 0, 2});
    p.fills.push_back(TermFill{term + 1, s_dom, 0, 0});
    for (uint32_t m = 0; m < L; m++) p.fills.push_back(TermFill{term + 2 + m, s_m + m, 0, 0});
    uint32_t fb = p.n_partials;
    p.comb(term, L + 2, CH);
    p.reduce_to(fb, p.n_partials - fb, pb);
    term += L + 2;
    // sum_U H_j m~_j (part of T2)
    uint32_t disc_begin = (uint32_t)p.disc_index.size();
    uint32_t und_begin = (uint32_t)p.und_scalar.size();
    uint32_t r = 0, u = 0;
    for (uint32_t i = 0; i < L; i++) {
      if (r < R && it.disclosed_indexes.data[r] == i) {
        p.disc_index.push_back(i);
        p.disc_scalar.push_back(s_m + i);
        r++;
      } else {
        p.term_gen[term + u] = 2 + i;
        p.fills.push_back(TermFill{term + u, rs + 5 + u, 0, 0});
        p.und_scalar.push_back(s_m + i);
        u++;
      }
    }
    // the two-term sums Bbar and T1 as multi-term products (shared doublings)
    uint32_t q = p.partials(5);
    p.var.push_back(VarJob{pb, kk + 0, q + 0});               // D = r2 B
    p.var.push_back(VarJob{pa, kk + 1, q + 1});               // Abar = r1 r2 A
    p.var_terms({{pb, kk + 2}, {pa, kk + 3}}, q + 2);         // Bbar = r1 r2 B - r1 r2 e A
    p.var_terms({{pa, kk + 4}, {pb, kk + 5}}, q + 3);         // T1 = r1 r2 e~ A + r2 r1~ B
    p.var.push_back(VarJob{pb, kk + 6, q + 4});               // T2 = r2 r3~ B + sum
    uint32_t ft2 = p.n_partials;
    p.comb(term, U, CH);
    term += U;
    p.groups2.push_back(ReduceGroup{q + 0, 1});
    p.out2.push_back(pd);
    p.groups2.push_back(ReduceGroup{q + 1, 1});
    p.out2.push_back(pab);
    p.groups2.push_back(ReduceGroup{q + 2, 1});
    p.out2.push_back(pbb);
    p.groups2.push_back(ReduceGroup{q + 3, 1});
    p.out2.push_back(pt1);
    // T2 partials: the last variable job's then the comb chunks [ft2, n_partials),
    // contiguous because ft2 == q + 5 (partials(5) was taken just before).
    p.groups2.push_back(ReduceGroup{q + 4, 1 + (p.n_partials - ft2)});
    p.out2.push_back(pt2);
    PgJob job{};
    job.ph_off = ph;
    job.proof_off = out;
    job.ph_len = (uint32_t)it.presentation_header.len;
    job.item = k;
    job.L = L;
    job.R = R;
    job.U = U;
    job.disc_begin = disc_begin;
    job.und_begin = und_begin;
    job.rs = rs;
    job.e_index = s_e;
    job.dom_index = s_dom;
    job.k_index = kk;
    job.a = pa;
    job.b = pb;
    job.d = pd;
    job.abar = pab;
    job.bbar = pbb;
    job.t1 = pt1;
    job.t2 = pt2;
    p.pg.push_back(job);
    b->out_off.push_back(out);
    b->out_len.push_back(272 + 32ull * U);
    out += 272 + 32ull * U;
  }
  p.out_bytes = out;
  if (mode != 2) {
    p.seeded = true;
    p.seeds_off = p.bytes(seeds.data(), seeds.size());
  }
  materialize(*b, p);
  return b;
}

// ------------------------------------------------------------------ launching
struct Timed {
  Batch& b;
  hipStream_t st;
  size_t idx = 0;
  Timed(Batch& bb, const char* name, hipStream_t on = nullptr) : b(bb), st(on ? on : bb.stream) {
    KernelTimer& t = b.timer;
    if (!t.on) return;
    auto it = std::find(t.names.begin(), t.names.end(), name);
    if (it == t.names.end()) {
      t.names.push_back(name);
      t.count.push_back(0);
      idx = t.names.size() - 1;
    } else {
      idx = it - t.names.begin();
    }
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0));
    check(hipEventCreate(&e1));
    t.starts.push_back(e0);
    t.stops.push_back(e1);
    t.event_name.push_back(idx);
    check(hipEventRecord(e0, st));
  }
  ~Timed() {
    if (!b.timer.on) return;
    (void)hipEventRecord(b.timer.stops.back(), st);
    b.timer.count[idx]++;
  }
};

void launch_batch(Batch& b) {
  if (b.nv == 0) return;
  hipStream_t s = b.stream;
  {
    Timed t(b, "fill_status");
    launch_fill_status(s, b.nv + b.n_agg_slots, b.status, b.gate, 1);
  }
  // Verify / ProofVerify: point decode and the variable-base products run on
  // the auxiliary stream while the message hashing, domain and comb MSM run
  // here; both need the decoded scalars and join before the reduction.
  const bool fork = (b.kind == CUBBS_HIP_BATCH_VERIFY || b.kind == CUBBS_HIP_BATCH_PROOF_VERIFY ||
                     b.kind == KIND_VERIFY_SET) &&
                    b.aux != nullptr;
  // Primitive programs decode without the subgroup check too.  The check
  // runs as extra lanes of the product launch, or, with a Pippenger MSM, as
  // its own launch on the auxiliary stream from the decode on, beside the
  // whole bucket pipeline, joined before the reduction.  The Pippenger
  // kernels raise their waves' priority so that the check's long lanes take
  // the issue slots they leave (2^16 points: 3.07-3.21 ms against 3.66-3.73
  // with the check beside the last stage only).
  const bool prim_split = b.kind == KIND_PRIM && b.n_dec_g1 && b.aux != nullptr;
  const bool side_check = prim_split && b.pip.terms;
  if (b.n_dec_fr) {
    Timed t(b, "decode_fr");
    launch_decode_fr(s, b.dec_fr, b.n_dec_fr, b.blob, b.scalars, b.status);
  }
  if (fork) {
    check(hipEventRecord(b.ev_fork, s));
    check(hipStreamWaitEvent(b.aux, b.ev_fork, 0));
    // the subgroup check of the decoded points runs as extra lanes of the
    // product kernel, beside the products rather than before them
    if (b.n_dec_g1) {
      Timed t(b, "decode_g1", b.aux);
      launch_decode_g1(b.aux, b.dec_g1, b.n_dec_g1, b.blob, b.points, b.pinf, b.status, false);
    }
    if (b.n_var || b.n_dec_g1 || b.n_var_sum) {
      Timed t(b, "var_mul_glv", b.aux);
      launch_var_mul(b.aux, b.var, b.n_var, b.points, b.pinf, b.scalars, b.partials, b.vtab, b.dec_g1,
                     b.n_dec_g1, b.status, b.var_sum, b.n_var_sum, b.sum_ws);
    }
    check(hipEventRecord(b.ev_join, b.aux));
  } else if (b.n_dec_g1) {
    {
      Timed t(b, "decode_g1");
      launch_decode_g1(s, b.dec_g1, b.n_dec_g1, b.blob, b.points, b.pinf, b.status, !prim_split);
    }
  }
  if (b.n_msg || b.n_dom) {
    Timed t(b, "hash_domain");
    launch_hash_stage(s, b.dom, b.n_dom, b.msg, b.n_msg, b.blob, b.sha_zpad, b.tail, b.tail_blocks,
                      b.pk_bytes, b.gen_comp, b.scalars);
  }
  if (b.kind == KIND_PROOF_GEN && b.seeded) {
    Timed t(b, "pg_chacha");
    launch_pg_chacha(s, b.pg, b.n_pg, b.seeds, b.scalars);
  }
  if (b.n_fills) {
    Timed t(b, "term_fill");
    launch_term_fill(s, b.fills, b.n_fills, b.scalars, b.sum_pairs);
  }
  if (b.kind == KIND_SIGN) {
    Timed t(b, "sign_e");
    launch_sign_e(s, b.sign, b.n_sign, b.blob, b.pk_status, b.scalars, b.status);
  }
  if (!fork && b.kind != KIND_PROOF_GEN && (b.n_var || b.n_var_sum || (prim_split && !side_check))) {
    const bool with_check = prim_split && !side_check;
    Timed t(b, "var_mul_glv");
    launch_var_mul(s, b.var, b.n_var, b.points, b.pinf, b.scalars, b.partials, b.vtab,
                   with_check ? b.dec_g1 : nullptr, with_check ? b.n_dec_g1 : 0, b.status, b.var_sum,
                   b.n_var_sum, b.sum_ws);
  }
  if (b.n_chunks) {
    Timed t(b, "comb_msm");
    launch_comb_msm(s, b.chunks, b.n_chunks, b.term_gen, b.scalars, b.tables, b.tables16, b.n16,
                    b.partials);
  }
  if (fork) check(hipStreamWaitEvent(s, b.ev_join, 0));
  if (b.pip.terms) {
    if (side_check) {
      check(hipEventRecord(b.ev_fork, s));
      check(hipStreamWaitEvent(b.aux, b.ev_fork, 0));
      Timed t(b, "subgroup", b.aux);
      launch_var_mul(b.aux, nullptr, 0, b.points, b.pinf, b.scalars, b.partials, b.vtab, b.dec_g1,
                     b.n_dec_g1, b.status);
    }
    {
      Timed t(b, "pip_sort");
      launch_fill_u32(s, b.pipb.cnt, b.pip.nbt + 1, 0);
      launch_pip_count(s, b.pip, b.pip_terms, b.pinf, b.scalars, b.pipb);
      launch_pip_scan(s, b.pip, b.pipb);
      launch_pip_scatter(s, b.pip, b.pip_terms, b.pinf, b.scalars, b.pipb);
    }
    {
      Timed t(b, "pip_accum");
      launch_pip_accum(s, b.pip, b.pip_terms, b.points, b.pipb);
    }
    {
      Timed t(b, "pip_bucket");
      launch_pip_bucket(s, b.pip, b.pipb);
    }
    {
      Timed t(b, "pip_halve");
      for (uint32_t st = 1; st <= b.pip.K; st++) launch_pip_halve(s, b.pip, st, b.pipb);
    }
    {
      Timed t(b, "pip_final");
      launch_pip_final(s, b.pip, b.pip_out, b.pipb, b.partials);
    }
    if (side_check) {
      check(hipEventRecord(b.ev_join, b.aux));
      check(hipStreamWaitEvent(s, b.ev_join, 0));
    }
  }
  for (const auto& lv : b.levels) {
    Timed t(b, "reduce_tree");
    launch_reduce_partials(s, lv.groups, lv.out, lv.n, b.partials);
  }
  if (b.n_groups1) {
    Timed t(b, "reduce");
    launch_reduce(s, b.groups1, b.out1, b.n_groups1, b.partials, b.points, b.pinf);
  }
  if (b.kind == KIND_PROOF_GEN) {
    {
      Timed t(b, "pg_scalars");
      launch_pg_scalars(s, b.pg, b.n_pg, b.scalars, b.status);
    }
    {
      Timed t(b, "var_mul_glv");
      launch_var_mul(s, b.var, b.n_var, b.points, b.pinf, b.scalars, b.partials, b.vtab, nullptr, 0,
                     nullptr, b.var_sum, b.n_var_sum, b.sum_ws);
    }
    {
      Timed t(b, "reduce2");
      launch_reduce(s, b.groups2, b.out2, b.n_groups2, b.partials, b.points, b.pinf);
    }
    {
      Timed t(b, "pg_finalize");
      launch_pg_finalize(s, b.pg, b.n_pg, b.blob, b.disc_index, b.disc_scalar, b.und_scalar,
                         b.points, b.pinf, b.scalars, b.status, b.out);
    }
  }
  if (b.kind == KIND_SIGN) {
    Timed t(b, "sign_finalize");
    launch_sign_finalize(s, b.sign, b.n_sign, b.points, b.pinf, b.scalars, b.status, b.out);
  }
  if (b.n_comp) {
    Timed t(b, "compress");
    launch_compress(s, b.comp, b.n_comp, b.points, b.pinf, b.out);
  }
  if (b.n_pvch) {
    Timed t(b, "pv_challenge");
    launch_pv_challenge(s, b.pvch, b.n_pvch, b.blob, b.disc_index, b.disc_scalar, b.points,
                        b.pinf, b.scalars, b.gate);
  }
  if (b.n_agg_slots) {
    // sums c_i a_i, c_i b_i of each aggregated group (ProofVerify; the
    // grouped Verify plan sums its groups in the main reduction)
    Timed t(b, "agg_reduce");
    launch_reduce(s, b.groups2, b.out2, b.n_groups2, b.partials, b.points, b.pinf);
  }
  if (b.n_pairs) {
    // Combined checks first, each level's results gating the next: the
    // supergroup checks, then the group checks (line blocks formed only for
    // groups of supergroups that failed), then the per-item checks (only for
    // items of groups that failed).
    uint32_t done = 0;
    const struct {
      uint32_t n;
      const char* name;
    } level[2] = {{b.n_spairs, "pairing_super"}, {b.n_gpairs, "pairing_groups"}};
    for (const auto& lv : level) {
      if (!lv.n) continue;
      {
        Timed t(b, "pair_lines");
        launch_pair_lines(s, b.pairs + done, lv.n, b.points, b.pinf, b.pk_status, b.pk_lines,
                          b.neg_g2_lines, b.l2buf + (size_t)done * PAIR_L2_WORDS,
                          done ? b.valid : nullptr);
      }
      Timed t(b, lv.name);
      launch_pairing(s, b.pairs + done, lv.n, b.points, b.pinf, b.pk_status,
                     b.l2buf + (size_t)done * PAIR_L2_WORDS, b.status, b.gate, b.valid);
      done += lv.n;
    }
    // the items of failing groups: their own products, comb terms and check
    // points, each launch over the list of its live jobs (the jobs of groups
    // whose check held would leave most waves partly idle, not empty)
    if (b.n_fb_sum || b.n_fb_chunks || b.n_fb_red) {
      Timed t(b, "gate_select");
      launch_fill_u32(s, b.gate_count, 3, 0);
      GateLists gl{};
      gl.gate[0] = b.fb_sum_gate;
      gl.gate[1] = b.fb_gate;
      gl.gate[2] = b.fb_red_gate;
      gl.n[0] = b.n_fb_sum;
      gl.n[1] = b.n_fb_chunks;
      gl.n[2] = b.n_fb_red;
      gl.list[0] = b.gate_list;
      gl.list[1] = b.gate_list + b.n_fb_sum;
      gl.list[2] = b.gate_list + b.n_fb_sum + b.n_fb_chunks;
      gl.count = b.gate_count;
      launch_gate_select(s, gl, b.valid);
    }
    if (b.n_fb_sum) {
      Timed t(b, "var_gated");
      launch_var_mul(s, nullptr, 0, b.points, b.pinf, b.scalars, b.partials, b.vtab, nullptr, 0,
                     nullptr, b.fb_sum, b.n_fb_sum, b.sum_ws, b.valid, b.gate_list, b.gate_count);
    }
    if (b.n_fb_chunks) {
      Timed t(b, "comb_gated");
      launch_comb_msm(s, b.fb_chunks, b.n_fb_chunks, b.term_gen, b.scalars, b.tables, b.tables16,
                      b.n16, b.partials, b.fb_gate, b.valid, b.gate_list + b.n_fb_sum, b.gate_count + 1);
    }
    if (b.n_fb_red) {
      Timed t(b, "reduce_gated");
      launch_reduce_gated(s, b.fb_red, b.n_fb_red, b.partials, b.points, b.pinf, b.valid,
                          b.gate_list + b.n_fb_sum + b.n_fb_chunks, b.gate_count + 2);
    }
    // after group checks, only the per-item checks that still need their own
    // pairing are launched, listed contiguously
    const uint32_t n_rest = b.n_pairs - done;
    uint32_t* list = done ? b.pair_list : nullptr;
    uint32_t* count = done ? b.pair_count : nullptr;
    if (done) {
      Timed t(b, "pair_select");
      launch_fill_u32(s, count, 1, 0);
      launch_pair_select(s, b.pairs + done, n_rest, b.pk_status, b.status, b.gate, b.valid, list, count);
    }
    {
      Timed t(b, "pair_lines");
      launch_pair_lines(s, b.pairs + done, n_rest, b.points, b.pinf, b.pk_status, b.pk_lines,
                        b.neg_g2_lines, b.l2buf + (size_t)done * PAIR_L2_WORDS, nullptr, list, count);
    }
    Timed t(b, "pairing");
    launch_pairing(s, b.pairs + done, n_rest, b.points, b.pinf, b.pk_status,
                   b.l2buf + (size_t)done * PAIR_L2_WORDS, b.status, b.gate, b.valid, list, count);
  }
  check(hipGetLastError());
}

void fetch_results(Batch& b) {
  // items, then the aggregated group checks (status and verdict slots)
  const uint32_t slots = b.nv + b.n_agg_slots;
  launch_copy(b.stream, b.h_res, b.status, slots);
  launch_copy(b.stream, b.h_res + b.res_stride, b.valid, slots);
  launch_copy(b.stream, b.h_res + 2 * b.res_stride, b.out, b.out_bytes);
  check(hipGetLastError());
  check(hipStreamSynchronize(b.stream));
  b.agg_passed = 0;
  for (uint32_t g = 0; g < b.n_agg_slots; g++) b.agg_passed += b.h_res[b.res_stride + b.nv + g] == 1;
  if (b.n_agg_slots) {
    // checks computed: the supergroup checks, and every later check whose
    // enclosing check did not hold
    uint64_t run = b.n_spairs;
    const uint32_t n_rest = b.n_pairs - b.n_spairs;
    for (uint32_t i = 0; i < n_rest; i++) {
      const uint32_t g = b.pair_group[i];
      run += g == PRIM_NO_SLOT || b.h_res[b.res_stride + g] != 1;
    }
    b.work.pair_checks_run = run;
    // the gated comb chunks and sums ran for the items of failing groups
    uint64_t terms = b.ungated_comb_terms, madds = b.ungated_comb_madds, adds = b.ungated_point_adds;
    uint64_t st[4] = {b.ungated_straus[0], b.ungated_straus[1], b.ungated_straus[2], b.ungated_straus[3]};
    for (const auto& g : b.gated_work) {
      if (b.h_res[b.res_stride + g.gate] == 1) continue;
      terms += g.terms;
      madds += g.madds;
      adds += g.adds;
      st[0] += g.sums;
      st[1] += g.sum_terms;
      st[2] += g.duals;
      st[3] += g.dual_terms;
    }
    b.work.comb_terms = terms;
    b.work.comb_madds = madds;
    b.work.point_adds = adds;
    b.work.straus_sums = st[0];
    b.work.straus_terms = st[1];
    b.work.straus_short_sums = st[2];
    b.work.straus_short_terms = st[3];
  }
}

// Copy results back and fill the caller's result array.  For Sign and
// ProofGen, outputs are written into the items' buffers (out_ptrs).
void collect_batch(Batch& b, CuBbsBbsGpuResult* results, uint8_t* const* out_ptrs,
                   size_t* out_lens) {
  b.h_status.resize(b.nv);
  b.h_valid.resize(b.nv);
  std::vector<uint8_t> hout;
  if (b.nv) {
    fetch_results(b);
    std::memcpy(b.h_status.data(), b.h_res, b.nv);
    std::memcpy(b.h_valid.data(), b.h_res + b.res_stride, b.nv);
    if (b.out_bytes) hout.assign(b.h_res + 2 * b.res_stride, b.h_res + 2 * b.res_stride + b.out_bytes);
  }
  for (size_t i = 0; i < b.n_total; i++) {
    results[i].bytes_written = 0;
    results[i].value = 0;
    if (b.pre_code[i] >= 0) {
      results[i].code = b.pre_code[i];
      results[i].value = b.pre_value[i];
      continue;
    }
    uint32_t k = b.dev_index[i];
    if (b.h_status[k] != 0) {
      results[i].code = CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      continue;
    }
    results[i].code = CUBBS_BBS_GPU_SUCCESS;
    if (b.kind == CUBBS_HIP_BATCH_VERIFY || b.kind == CUBBS_HIP_BATCH_PROOF_VERIFY ||
        b.kind == KIND_VERIFY_SET) {
      results[i].value = b.h_valid[k];
    } else if (b.kind == KIND_SIGN) {
      if (out_ptrs) std::memcpy(out_ptrs[i], hout.data() + 80ull * k, 80);
      if (out_lens) out_lens[i] = 80;
      results[i].bytes_written = 80;
    } else if (b.kind == KIND_PROOF_GEN) {
      if (out_ptrs) std::memcpy(out_ptrs[i], hout.data() + b.out_off[k], b.out_len[k]);
      if (out_lens) out_lens[i] = b.out_len[k];
      results[i].bytes_written = b.out_len[k];
    }
  }
}

void capture_batch(Batch& b) {
  if (b.graph) return;
  hipStream_t cs = nullptr, ca = nullptr;
  check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamCreateWithFlags(&ca, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(cs);
    check(e);
  }
  const hipStream_t s0 = b.stream, a0 = b.aux;
  const bool timing = b.timer.on;
  b.stream = cs;
  b.aux = a0 ? ca : nullptr;
  b.timer.on = false;
  hipGraph_t g = nullptr;
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    try {
      launch_batch(b);
    } catch (const GpuError&) {
      e = hipErrorUnknown;
    }
    hipError_t e2 = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) e = e2;
  }
  b.stream = s0;
  b.aux = a0;
  b.timer.on = timing;
  if (e == hipSuccess) e = hipGraphInstantiate(&b.graph, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  (void)hipStreamDestroy(cs);
  (void)hipStreamDestroy(ca);
  if (e != hipSuccess) b.graph = nullptr;
  check(e);
}

Batch::~Batch() {
  if (stream) (void)hipStreamSynchronize(stream);
  if (graph) (void)hipGraphExecDestroy(graph);
  for (size_t i = 0; i < timer.starts.size(); i++) {
    (void)hipEventDestroy(timer.starts[i]);
    (void)hipEventDestroy(timer.stops[i]);
  }
  if (aux) (void)hipStreamSynchronize(aux);
  for (void* p : owned) (void)hipFree(p);
  if (hold.held) {
    Context& c = ctx();
    std::lock_guard<std::mutex> g(c.mu);
    release_locked(c, hold);
  }
  if (from_arena) return;  // the arena owns memory and events
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (h_in) (void)hipHostFree(h_in);
  if (h_res) (void)hipHostFree(h_res);
  if (d_in) (void)hipFree(d_in);
  if (d_work) (void)hipFree(d_work);  // streams belong to the context's pool
}

}  // namespace cubbs
