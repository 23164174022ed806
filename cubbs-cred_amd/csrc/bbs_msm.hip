// Pippenger bucket MSM for large variable-base multi-scalar multiplications
// (the reference's large-MSM route, route id 3:
// native/cuda/src/msm_dispatch.cuh:110-155).
//
// Each scalar k is split k = k1 + k2 * lambda (GLV, two 128-bit halves) and
// each half recoded into W signed c-bit digits.  Every non-zero digit d of
// window w sends +-P (half 0) or +-phi(P) (half 1) to bucket (w, |d|).  The
// (term, half, window) entries are sorted by bucket with a counting sort: an
// atomic histogram, one scan, an atomic scatter.  Buckets are cut into
// pieces of at most S entries, one lane per piece sums its entries with mixed
// additions, and one lane per bucket adds its pieces.
//
// Per window, sum_{b=1..B} b * S_b is formed without a B-long running sum.
// With X_i = S_{i+1} (i = 0..B-1), K = log2 B in-place steps do, at step s
// (M = B >> s), X_i += X_{i+M} for i in [0, M) and in [B >> u, (B >> u) + M)
// for u = 1..s-1.  Afterwards X_0 = sum_i X_i and X_{2^j} is the sum of the
// range whose weight is 2^j, so sum_i i * X_i = sum_j 2^j X_{2^j}: 2B adds in
// K parallel steps.  A Horner pass over j gives each window sum, a Horner
// pass over the windows (c doublings apart) the MSM.  The order of entries
// inside a bucket depends on the scatter's atomics, the group sums do not,
// and the caller converts the result to affine form.
#include <algorithm>
#include <cstdlib>

#include "bbs_internal.h"
#include "dev_io.hpp"

namespace cubbs {

#define GRID(n, b) dim3(((n) + (b) - 1) / (b)), dim3(b)

namespace {

// Signed c-bit digits of the two 128-bit GLV halves of k: calls
// f(half, window, bucket = |digit| - 1, negative) for each non-zero digit.
// Digits lie in (-2^(c-1), 2^(c-1)]; W c >= 129 leaves no final carry.
template <class F>
DEV void pip_digits(const uint32_t k[8], uint32_t c, uint32_t W, F f) {
  uint32_t k1[4], k2[4];
  glv_split(k1, k2, k);
  const uint64_t mask = (1ull << c) - 1;
  const uint32_t half = 1u << (c - 1), full = 1u << c;
#pragma unroll
  for (uint32_t h = 0; h < 2; h++) {
    uint64_t lo = h ? (k2[0] | (uint64_t)k2[1] << 32) : (k1[0] | (uint64_t)k1[1] << 32);
    uint64_t hi = h ? (k2[2] | (uint64_t)k2[3] << 32) : (k1[2] | (uint64_t)k1[3] << 32);
    uint32_t carry = 0;
    for (uint32_t w = 0; w < W; w++) {
      const uint32_t d = (uint32_t)(lo & mask) + carry;
      lo = (lo >> c) | (hi << (64 - c));
      hi >>= c;
      uint32_t mag = d, neg = 0;
      carry = 0;
      if (d > half) {
        mag = full - d;
        neg = 1;
        carry = 1;
      }
      if (mag) f(h, w, mag - 1, neg);
    }
  }
}

}  // namespace

// FpM model: 2n entries per window at ~12 FpM (mixed add + piece share),
// ~2 full additions per bucket for the halving and the bucket sums.
uint32_t pip_model_window(double n) {
  uint32_t best = 8;
  double best_cost = 0;
  for (uint32_t c = 4; c <= 16; c++) {
    const uint32_t W = (129 + c - 1) / c;
    const double cost = W * (24.0 * n + 32.0 * (double)(1u << (c - 1)));
    if (c == 4 || cost < best_cost) {
      best = c;
      best_cost = cost;
    }
  }
  return best;
}

PipShape pip_shape(uint32_t terms, uint32_t jobs) {
  PipShape sh{};
  jobs = std::max<uint32_t>(jobs, 1);
  const double n = (double)terms / jobs;
  const uint32_t best = pip_model_window(n);
  // precedence: the calling thread's forced shape, the process setting, the
  // tuned table, the model above
  uint32_t c = best, S = 16, tc, ts;
  if (tuned_pip((uint32_t)n, tc, ts)) {
    c = tc;
    S = ts;
  }
  if (knobs().pip_window) c = knobs().pip_window;
  if (knobs().pip_piece) S = knobs().pip_piece;
  if (tune_force().pip_c) c = tune_force().pip_c;
  if (tune_force().pip_s) S = tune_force().pip_s;
  sh.c = std::min<uint32_t>(16, std::max<uint32_t>(4, c));
  sh.W = (129 + sh.c - 1) / sh.c;
  sh.K = sh.c - 1;
  sh.B = 1u << sh.K;
  sh.S = std::min<uint32_t>(256, std::max<uint32_t>(1, S));
  sh.jobs = jobs;
  sh.terms = terms;
  sh.nbt = jobs * sh.W * sh.B;
  sh.max_entries = 2 * terms * sh.W;
  sh.max_pieces = sh.max_entries / sh.S + sh.nbt + 1;
  return sh;
}

// ------------------------------------------------------------------ fill
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t n, uint32_t v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, GRID(n, 64), 0, s, p, n, v);
}

// ------------------------------------------------------------------ histogram
__global__ __launch_bounds__(256) void k_pip_count(PipShape sh, const PipTerm* __restrict__ terms,
                                                  const uint8_t* __restrict__ pinf,
                                                  const uint32_t* __restrict__ scalars,
                                                  uint32_t* __restrict__ cnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sh.terms) return;
  const PipTerm pt = terms[t];
  if (pinf[pt.point]) return;
  Fr k;
  load_fr(k, scalars + (size_t)pt.scalar * FR_WORDS);
  const uint32_t base = pt.job * sh.W * sh.B;
  pip_digits(k.v, sh.c, sh.W, [&](uint32_t, uint32_t w, uint32_t b, uint32_t) {
    atomicAdd(&cnt[base + w * sh.B + b], 1u);
  });
}
void launch_pip_count(hipStream_t s, const PipShape& sh, const PipTerm* terms, const uint8_t* pinf,
                      const uint32_t* scalars, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_count, GRID(sh.terms, 64), 0, s, sh, terms, pinf, scalars, bf.cnt);
}

// ------------------------------------------------------------------ scan
// One block: exclusive prefix sums of the bucket counts (entry offsets) and
// of ceil(count / S) (piece offsets); writes the piece -> bucket map.
constexpr int SCAN_THREADS = 1024;
__global__ __launch_bounds__(SCAN_THREADS) void k_pip_scan(PipShape sh, PipBuffers bf) {
  __shared__ uint32_t sa[SCAN_THREADS], sb[SCAN_THREADS];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (sh.nbt + SCAN_THREADS - 1) / SCAN_THREADS;
  const uint32_t lo = min(sh.nbt, tid * per), hi = min(sh.nbt, lo + per);
  uint32_t a = 0, p = 0;
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t c = bf.cnt[i];
    a += c;
    p += (c + sh.S - 1) / sh.S;
  }
  sa[tid] = a;
  sb[tid] = p;
  __syncthreads();
  for (uint32_t d = 1; d < SCAN_THREADS; d <<= 1) {
    const uint32_t va = tid >= d ? sa[tid - d] : 0, vb = tid >= d ? sb[tid - d] : 0;
    __syncthreads();
    sa[tid] += va;
    sb[tid] += vb;
    __syncthreads();
  }
  uint32_t ea = sa[tid] - a, eb = sb[tid] - p;
  for (uint32_t i = lo; i < hi; i++) {
    const uint32_t c = bf.cnt[i];
    const uint32_t np = (c + sh.S - 1) / sh.S;
    bf.off[i] = ea;
    bf.cursor[i] = ea;
    bf.poff[i] = eb;
    for (uint32_t k = 0; k < np && eb + k < sh.max_pieces; k++) bf.pmap[eb + k] = i;
    ea += c;
    eb += np;
  }
  if (tid == SCAN_THREADS - 1) {
    const uint32_t pieces = min(sb[SCAN_THREADS - 1], sh.max_pieces);
    bf.off[sh.nbt] = sa[SCAN_THREADS - 1];
    bf.poff[sh.nbt] = pieces;
    bf.hdr[0] = pieces;
    bf.hdr[1] = sa[SCAN_THREADS - 1];
  }
}
void launch_pip_scan(hipStream_t s, const PipShape& sh, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_scan, dim3(1), dim3(SCAN_THREADS), 0, s, sh, bf);
}

// ------------------------------------------------------------------ scatter
__global__ __launch_bounds__(256) void k_pip_scatter(PipShape sh, const PipTerm* __restrict__ terms,
                                                    const uint8_t* __restrict__ pinf,
                                                    const uint32_t* __restrict__ scalars,
                                                    uint32_t* __restrict__ cursor,
                                                    uint32_t* __restrict__ sorted) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= sh.terms) return;
  const PipTerm pt = terms[t];
  if (pinf[pt.point]) return;
  Fr k;
  load_fr(k, scalars + (size_t)pt.scalar * FR_WORDS);
  const uint32_t base = pt.job * sh.W * sh.B;
  pip_digits(k.v, sh.c, sh.W, [&](uint32_t h, uint32_t w, uint32_t b, uint32_t neg) {
    const uint32_t pos = atomicAdd(&cursor[base + w * sh.B + b], 1u);
    if (pos < sh.max_entries) sorted[pos] = t << 2 | h << 1 | neg;
  });
}
void launch_pip_scatter(hipStream_t s, const PipShape& sh, const PipTerm* terms,
                        const uint8_t* pinf, const uint32_t* scalars, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_scatter, GRID(sh.terms, 64), 0, s, sh, terms, pinf, scalars, bf.cursor,
                     bf.sorted);
}

// ------------------------------------------------------------------ piece sums
__global__ __launch_bounds__(64) void k_pip_accum(PipShape sh, const PipTerm* __restrict__ terms,
                                                 const uint32_t* __restrict__ points, PipBuffers bf) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= sh.max_pieces || p >= bf.hdr[0]) return;
  const uint32_t b = bf.pmap[p];
  const uint32_t first = bf.off[b] + (p - bf.poff[b]) * sh.S;
  const uint32_t end = min(first + sh.S, bf.off[b] + bf.cnt[b]);
  Fp beta;
  fp_load_const(beta, bls::GLV_BETA);
  G1J acc;
  pt_set_inf(acc);
  for (uint32_t e = first; e < end; e++) {
    const uint32_t v = bf.sorted[e];
    G1A q;
    load_g1a(q, points + (size_t)terms[v >> 2].point * G1A_WORDS);
    if (v & 2) fp_mul(q.x, q.x, beta);
    if (v & 1) fp_neg(q.y, q.y);
    g1_madd_in(acc, q);
  }
  store_g1j(bf.pieces + (size_t)p * G1J_WORDS, acc);
}
void launch_pip_accum(hipStream_t s, const PipShape& sh, const PipTerm* terms,
                      const uint32_t* points, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_accum, GRID(sh.max_pieces, 64), 0, s, sh, terms, points, bf);
}

// ------------------------------------------------------------------ bucket sums
__global__ __launch_bounds__(64) void k_pip_bucket(PipShape sh, PipBuffers bf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sh.nbt) return;
  G1J acc;
  pt_set_inf(acc);
  const uint32_t end = min(bf.poff[i + 1], sh.max_pieces);
  for (uint32_t p = bf.poff[i]; p < end; p++) {
    G1J x;
    load_g1j(x, bf.pieces + (size_t)p * G1J_WORDS);
    g1_add_in(acc, x);
  }
  store_g1j(bf.buckets + (size_t)i * G1J_WORDS, acc);
}
void launch_pip_bucket(hipStream_t s, const PipShape& sh, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_bucket, GRID(sh.nbt, 64), 0, s, sh, bf);
}

// ------------------------------------------------------------------ halving step
// Step s (1..K), M = B >> s: X_i += X_{i+M} for i in [0, M) and, for
// u = 1..s-1, in [B >> u, (B >> u) + M); reads and writes of one step are
// disjoint (the ranges written start at multiples of 2M, those read at an
// odd multiple of M).
__global__ __launch_bounds__(64) void k_pip_halve(PipShape sh, uint32_t step, uint32_t M,
                                                 uint32_t n, uint32_t* __restrict__ X) {
  const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const uint32_t per = step * M;
  const uint32_t jw = idx / per, q = idx % per;
  const uint32_t u = q / M, r = q % M;
  const uint32_t i = (u ? (sh.B >> u) : 0u) + r;
  uint32_t* base = X + (size_t)jw * sh.B * G1J_WORDS;
  G1J a, b;
  load_g1j(a, base + (size_t)i * G1J_WORDS);
  load_g1j(b, base + (size_t)(i + M) * G1J_WORDS);
  g1_add_in(a, b);
  store_g1j(base + (size_t)i * G1J_WORDS, a);
}
void launch_pip_halve(hipStream_t s, const PipShape& sh, uint32_t step, const PipBuffers& bf) {
  if (!sh.terms || step < 1 || step > sh.K) return;
  const uint32_t M = sh.B >> step;
  const uint32_t n = sh.jobs * sh.W * step * M;
  hipLaunchKernelGGL(k_pip_halve, GRID(n, 64), 0, s, sh, step, M, n, bf.buckets);
}

// ------------------------------------------------------------------ window combination
// One block per MSM: lane w forms window w's sum X_0 + sum_j 2^j X_{2^j};
// lane 0 then combines the windows, c doublings apart.
__global__ __launch_bounds__(64) void k_pip_final(PipShape sh, const uint32_t* __restrict__ job_out,
                                                 const uint32_t* __restrict__ X,
                                                 uint32_t* __restrict__ partials) {
  __shared__ uint32_t win[64 * G1J_WORDS];
  const uint32_t job = blockIdx.x, w = threadIdx.x;
  if (w < sh.W) {
    const uint32_t* base = X + (size_t)(job * sh.W + w) * sh.B * G1J_WORDS;
    G1J acc, x;
    pt_set_inf(acc);
    for (int j = (int)sh.K - 1; j >= 0; j--) {
      g1_dbl_in(acc);
      load_g1j(x, base + (size_t)(1u << j) * G1J_WORDS);
      g1_add_in(acc, x);
    }
    load_g1j(x, base);
    g1_add_in(acc, x);
    store_g1j(win + w * G1J_WORDS, acc);
  }
  __syncthreads();
  if (w == 0) {
    G1J r, x;
    load_g1j(r, win + (sh.W - 1) * G1J_WORDS);
    for (int v = (int)sh.W - 2; v >= 0; v--) {
      for (uint32_t d = 0; d < sh.c; d++) g1_dbl_in(r);
      load_g1j(x, win + v * G1J_WORDS);
      g1_add_in(r, x);
    }
    store_g1j(partials + (size_t)job_out[job] * G1J_WORDS, r);
  }
}
void launch_pip_final(hipStream_t s, const PipShape& sh, const uint32_t* job_out,
                      const PipBuffers& bf, uint32_t* partials) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_final, dim3(sh.jobs), dim3(64), 0, s, sh, job_out, bf.buckets, partials);
}

}  // namespace cubbs
