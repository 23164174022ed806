// Measured autotuner for the MSM launch shapes.
//
// For each size class it stages a synthetic batch per candidate shape, times a
// few launches with HIP events on the batch's stream and keeps the fastest for
// the rest of the process:
//  * Pippenger MSMs of n = 2^k terms (k = 10 .. max): window widths c within
//    two of the cost model's choice, pieces of S in {8, 16, 32} entries;
//  * comb MSMs: 32-term rows over the ciphersuite generators, batches of 2^k
//    fixed-base terms (k = 12, 14, 16, 18): 1, 2, 4, 8 or 16 terms per lane.
// Lookups use the nearest tuned class within two powers of two.  The table is
// reported as text (one line per class) and can be installed again at start-up
// with cubbs_hip_tune_load.  The table also takes lines for the whole-operation
// Verify / ProofVerify pipelines (comb terms per lane and aggregation group
// size by items per batch), which a caller measures by running sets of such
// batches in flight, as it runs them, and loads the same way.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_batch.h"
#include "host_ctx.h"

namespace cubbs {
namespace {

constexpr int CLASSES = 41;

struct Table {
  std::mutex mu;
  uint8_t pip_c[CLASSES] = {};
  uint16_t pip_s[CLASSES] = {};
  uint8_t comb[CLASSES] = {};
  // whole-operation pipelines, [kind][log2 items]: comb chunk, group size
  uint8_t pipe_chunk[2][CLASSES] = {};
  uint8_t pipe_group[2][CLASSES] = {};
  uint8_t pipe_super[2][CLASSES] = {};
  uint8_t pipe_straus[2][CLASSES] = {};
};
Table& table() {
  static Table t;
  return t;
}

// nearest power of two (as an exponent), capped to the table
uint32_t log2_near(uint64_t n) {
  if (n <= 1) return 0;
  uint32_t k = 63 - __builtin_clzll(n);
  if (k < 63 && (n - (1ull << k)) * 2 >= (1ull << k)) k++;  // round to the nearer power
  return std::min<uint32_t>(k, CLASSES - 1);
}

template <class Has>
int nearest_class(uint32_t k, Has has) {
  for (int d = 0; d <= 2; d++)
    for (int sg : {-1, 1}) {
      const int kk = (int)k + sg * d;
      if (kk >= 0 && kk < CLASSES && has(kk)) return kk;
    }
  return -1;
}

// Average milliseconds of `iters` launches of a staged batch (after one
// warm-up launch).
double time_batch(Batch& b, int iters) {
  launch_batch(b);
  check(hipStreamSynchronize(b.stream));
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0));
  check(hipEventCreate(&e1));
  check(hipEventRecord(e0, b.stream));
  for (int i = 0; i < iters; i++) launch_batch(b);
  check(hipEventRecord(e1, b.stream));
  check(hipEventSynchronize(e1));
  float ms = 0;
  check(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / iters;
}

// random scalar below 2^254 < r, big-endian
void rnd_scalar(std::mt19937_64& rng, uint8_t* out) {
  for (int i = 0; i < 4; i++) {
    uint64_t v = rng();
    for (int j = 0; j < 8; j++) out[8 * i + j] = (uint8_t)(v >> (8 * (7 - j)));
  }
  out[0] &= 0x3f;
}

// compressed G1 generator
const uint8_t G1_GEN[48] = {
    0x97, 0xf1, 0xd3, 0xa7, 0x31, 0x97, 0xd7, 0x94, 0x26, 0x95, 0x63, 0x8c, 0x4f, 0xa9, 0xac, 0x0f,
    0xc3, 0x68, 0x8c, 0x4f, 0x97, 0x74, 0xb9, 0x05, 0xa1, 0x4e, 0x3a, 0x3f, 0x17, 0x1b, 0xac, 0x58,
    0x6c, 0x55, 0xe8, 0x3f, 0xf9, 0x7a, 0x1a, 0xef, 0xfb, 0x3a, 0xf0, 0x0a, 0xdb, 0x22, 0xc6, 0xbb};

struct ForceGuard {
  ForceGuard(uint32_t c, uint32_t s, uint32_t chunk) { tune_force() = TuneForce{c, s, chunk}; }
  ~ForceGuard() { tune_force() = TuneForce{}; }
};

}  // namespace

TuneForce& tune_force() {
  thread_local TuneForce f;
  return f;
}

bool tuned_pip(uint32_t n, uint32_t& c, uint32_t& S) {
  Table& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  const int k = nearest_class(log2_near(n), [&](int kk) { return t.pip_c[kk] != 0; });
  if (k < 0) return false;
  c = t.pip_c[k];
  S = t.pip_s[k];
  return true;
}

uint32_t tuned_comb_chunk(uint64_t terms) {
  Table& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  const int k = nearest_class(log2_near(terms), [&](int kk) { return t.comb[kk] != 0; });
  return k < 0 ? 0 : t.comb[k];
}

bool tuned_pipeline(int kind, uint64_t items, uint32_t& comb_chunk, uint32_t& agg_group,
                    uint32_t& agg_super, uint32_t* agg_straus) {
  if (kind < 0 || kind > 1) return false;
  Table& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  const int k = nearest_class(log2_near(items), [&](int kk) { return t.pipe_chunk[kind][kk] != 0; });
  if (k < 0) return false;
  comb_chunk = t.pipe_chunk[kind][k];
  agg_group = t.pipe_group[kind][k];
  agg_super = t.pipe_super[kind][k];
  if (agg_straus) *agg_straus = t.pipe_straus[kind][k];
  return true;
}

}  // namespace cubbs

using namespace cubbs;

extern "C" {

int32_t cubbs_hip_autotune(uint32_t max_log2_terms, char* report, size_t report_cap) {
  if (max_log2_terms < 10 || max_log2_terms > 20) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  std::string rep;
  try {
    std::mt19937_64 rng(0x5eed);
    // 4096 distinct points k_i G1, reused cyclically by the larger MSMs
    const size_t NP = 4096;
    std::vector<uint8_t> base(48 * NP), ks(32 * NP), pts(48 * NP), ok(NP);
    for (size_t i = 0; i < NP; i++) {
      std::memcpy(&base[48 * i], G1_GEN, 48);
      rnd_scalar(rng, &ks[32 * i]);
    }
    int32_t rc = cubbs_hip_test_g1_mul(base.data(), ks.data(), NP, pts.data(), ok.data());
    if (rc) return rc;
    char line[160];
    // ---- Pippenger classes
    for (uint32_t k = 10; k <= max_log2_terms; k++) {
      const size_t n = (size_t)1 << k;
      std::vector<uint8_t> sc(32 * n);
      for (size_t i = 0; i < n; i++) rnd_scalar(rng, &sc[32 * i]);
      std::vector<CuBbsBbsGpuByteSlice> ps(n), ss(n);
      for (size_t i = 0; i < n; i++) {
        ps[i] = CuBbsBbsGpuByteSlice{&pts[48 * (i % NP)], 48};
        ss[i] = CuBbsBbsGpuByteSlice{&sc[32 * i], 32};
      }
      CuBbsBbsGpuG1MsmItem item{};
      item.points = ps.data();
      item.scalars = ss.data();
      item.term_count = n;
      const uint32_t cm = pip_model_window((double)n);
      double best = 1e30;
      uint32_t bc = cm, bs = 16;
      for (uint32_t c = std::max<uint32_t>(4, cm - 2); c <= std::min<uint32_t>(16, cm + 2); c++)
        for (uint32_t S : {8u, 16u, 32u}) {
          ForceGuard fg(c, S, 0);
          int32_t st = 0;
          auto b = stage_primitive(CUBBS_HIP_BATCH_G1_MSM, &item, 1, st);
          if (!b) return st ? st : CUBBS_BBS_GPU_ERR_INTERNAL;
          const double ms = time_batch(*b, 3);
          if (ms < best) {
            best = ms;
            bc = c;
            bs = S;
          }
        }
      {
        Table& t = table();
        std::lock_guard<std::mutex> g(t.mu);
        t.pip_c[k] = (uint8_t)bc;
        t.pip_s[k] = (uint16_t)bs;
      }
      std::snprintf(line, sizeof line, "pippenger log2_terms=%u c=%u S=%u ms=%.3f\n", k, bc, bs, best);
      rep += line;
    }
    // ---- comb classes: rows of 32 terms over [P1, Q1, H_1 .. H_30]
    std::vector<uint8_t> gens(48 * 32);
    rc = cubbs_hip_generators(32, gens.data());
    if (rc) return rc;
    std::vector<CuBbsBbsGpuByteSlice> gp(32);
    for (int i = 0; i < 32; i++) gp[i] = CuBbsBbsGpuByteSlice{&gens[48 * i], 48};
    for (uint32_t k : {12u, 14u, 16u, 18u}) {
      const size_t rows = ((size_t)1 << k) / 32;
      std::vector<uint8_t> sc(32 * 32 * rows);
      for (size_t i = 0; i < 32 * rows; i++) rnd_scalar(rng, &sc[32 * i]);
      CuBbsBbsGpuG1MsmCommonPointsBatchItem grp{};
      grp.points = gp.data();
      grp.term_count = 32;
      grp.scalar_rows_be = sc.data();
      grp.scalar_row_stride = 32 * 32;
      grp.output_count = rows;
      double best = 1e30;
      uint32_t bch = 4;
      for (uint32_t ch : {1u, 2u, 4u, 8u, 16u}) {
        ForceGuard fg(0, 0, ch);
        int32_t st = 0;
        auto b = stage_primitive(CUBBS_HIP_BATCH_G1_MSM_COMMON_POINTS, &grp, 1, st);
        if (!b) return st ? st : CUBBS_BBS_GPU_ERR_INTERNAL;
        const double ms = time_batch(*b, 3);
        if (ms < best) {
          best = ms;
          bch = ch;
        }
      }
      {
        Table& t = table();
        std::lock_guard<std::mutex> g(t.mu);
        t.comb[k] = (uint8_t)bch;
      }
      std::snprintf(line, sizeof line, "comb log2_terms=%u chunk=%u ms=%.3f\n", k, bch, best);
      rep += line;
    }
  } catch (...) {  // a GPU error, or host memory exhausted
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  if (report && report_cap) {
    const size_t n = std::min(rep.size(), report_cap - 1);
    std::memcpy(report, rep.data(), n);
    report[n] = 0;
  }
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_test_force_shapes(uint32_t pip_c, uint32_t pip_s, uint32_t comb_chunk) {
  if ((pip_c && (pip_c < 4 || pip_c > 16)) || pip_s > 256 || comb_chunk > 64)
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  tune_force() = TuneForce{pip_c, pip_s, comb_chunk};
  return CUBBS_BBS_GPU_SUCCESS;
}

int32_t cubbs_hip_tune_load(const char* text) {
  if (!text) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  Table& t = table();
  std::lock_guard<std::mutex> g(t.mu);
  if (!*text) {  // empty table: back to the cost model and defaults
    std::memset(t.pip_c, 0, sizeof t.pip_c);
    std::memset(t.pip_s, 0, sizeof t.pip_s);
    std::memset(t.comb, 0, sizeof t.comb);
    std::memset(t.pipe_chunk, 0, sizeof t.pipe_chunk);
    std::memset(t.pipe_group, 0, sizeof t.pipe_group);
    std::memset(t.pipe_super, 0, sizeof t.pipe_super);
    std::memset(t.pipe_straus, 0, sizeof t.pipe_straus);
    return CUBBS_BBS_GPU_SUCCESS;
  }
  const char* p = text;
  while (*p) {
    const char* e = std::strchr(p, '\n');
    const std::string ln(p, e ? (size_t)(e - p) : std::strlen(p));
    unsigned k = 0, a = 0, b = 0;
    char kind[16] = {};
    if (std::sscanf(ln.c_str(), "pippenger log2_terms=%u c=%u S=%u", &k, &a, &b) == 3) {
      if (k >= CLASSES || a < 4 || a > 16 || b < 1 || b > 256) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      t.pip_c[k] = (uint8_t)a;
      t.pip_s[k] = (uint16_t)b;
    } else if (std::sscanf(ln.c_str(), "comb log2_terms=%u chunk=%u", &k, &a) == 2) {
      if (k >= CLASSES || a < 1 || a > 64) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      t.comb[k] = (uint8_t)a;
    } else if (std::sscanf(ln.c_str(), "pipeline kind=%15s log2_items=%u chunk=%u group=%u", kind, &k, &a,
                           &b) == 4) {
      // optional: " super=M" (groups per supergroup, 0 or 2..8) and
      // " straus=S" (signatures per product job of a Verify group, 0..4)
      unsigned sup = 0, st = 0;
      const char* sp = std::strstr(ln.c_str(), " super=");
      if (sp && std::sscanf(sp, " super=%u", &sup) != 1) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      const char* ss = std::strstr(ln.c_str(), " straus=");
      if (ss && std::sscanf(ss, " straus=%u", &st) != 1) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      const int kd = std::strcmp(kind, "verify") == 0 ? 0 : std::strcmp(kind, "proof_verify") == 0 ? 1 : -1;
      if (kd < 0 || k >= CLASSES || a < 1 || a > 64 || b < 2 || b > 8 || sup == 1 || sup > 8 ||
          st > VAR_TERMS_MAX)
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      t.pipe_chunk[kd][k] = (uint8_t)a;
      t.pipe_group[kd][k] = (uint8_t)b;
      t.pipe_super[kd][k] = (uint8_t)sup;
      t.pipe_straus[kd][k] = (uint8_t)st;
    } else if (!ln.empty()) {
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    }
    if (!e) break;
    p = e + 1;
  }
  return CUBBS_BBS_GPU_SUCCESS;
}

}  // extern "C"
