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
// additions (in XYZZ coordinates), and one lane per bucket adds its pieces.
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
#include <stdexcept>

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
// The Pippenger stages raise their waves' issue priority above the product
// kernel's: the prime-order subgroup checks of the MSM's points run beside
// them (a launch of long, VALU-bound lanes on another stream), and at equal
// priority they took the issue slots the bucket accumulation needs
// (0.38 -> 1.3 ms).
DEV void pip_priority() { wave_priority<PRIO_PIP>(); }

__global__ __launch_bounds__(256) void k_pip_count(PipShape sh, const PipTerm* __restrict__ terms,
                                                  const uint8_t* __restrict__ pinf,
                                                  const uint32_t* __restrict__ scalars,
                                                  uint32_t* __restrict__ cnt) {
  pip_priority();
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
  pip_priority();
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
  pip_priority();
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
  pip_priority();
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= sh.max_pieces || p >= bf.hdr[0]) return;
  const uint32_t b = bf.pmap[p];
  const uint32_t first = bf.off[b] + (p - bf.poff[b]) * sh.S;
  const uint32_t end = min(first + sh.S, bf.off[b] + bf.cnt[b]);
  Fp beta;
  fp_load_const(beta, bls::GLV_BETA);
  G1Z acc;  // XYZZ: the cheaper mixed addition
  g1z_set_inf(acc);
  for (uint32_t e = first; e < end; e++) {
    const uint32_t v = bf.sorted[e];
    G1A q;
    load_g1a(q, points + (size_t)terms[v >> 2].point * G1A_WORDS);
    if (v & 2) fp_mul(q.x, q.x, beta);
    fp_cond_neg_lazy(q.y, (v & 1) != 0);
    g1z_madd(acc, q);
  }
  G1J out;
  g1z_to_jacobian(out, acc);
  store_g1j(bf.pieces + (size_t)p * G1J_WORDS, out);
}
void launch_pip_accum(hipStream_t s, const PipShape& sh, const PipTerm* terms,
                      const uint32_t* points, const PipBuffers& bf) {
  if (!sh.terms) return;
  hipLaunchKernelGGL(k_pip_accum, GRID(sh.max_pieces, 64), 0, s, sh, terms, points, bf);
}

// ------------------------------------------------------------------ bucket sums
__global__ __launch_bounds__(64) void k_pip_bucket(PipShape sh, PipBuffers bf) {
  pip_priority();
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

// ------------------------------------------------------------------ window combination
// The window sums and their combination are chains of dependent group
// operations (about 129 doublings in all), so their latency, not their work,
// sets the time.  Each chain runs on the four lanes of a quad: the
// independent field products of one doubling or addition run side by side,
// one per lane, and DPP quad broadcasts (quad_perm [j, j, j, j]) hand every
// product to all four lanes.  A doubling takes 3 product latencies instead
// of 7, an addition 5 instead of 16.  The four lanes of a quad always hold
// the same point values, so every branch is uniform within the quad.

template <int J>
DEV void fp_quad_bcast(Fp& r, const Fp& a) {
#pragma unroll
  for (int k = 0; k < FPW; k++)
    r.v[k] = (uint32_t)__builtin_amdgcn_mov_dpp((int)a.v[k], J * 0x55, 0xf, 0xf, false);
}
// r = (a0, a1, a2, a3)[q] by two levels of selects on the bits of q (a
// chain of comparisons with q would be turned into an indexed table in
// scratch memory)
DEV void fp_sel4(Fp& r, uint32_t q, const Fp& a0, const Fp& a1, const Fp& a2, const Fp& a3) {
  const bool b0 = (q & 1) != 0, b1 = (q & 2) != 0;
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    const uint32_t lo = b0 ? a1.v[k] : a0.v[k], hi = b0 ? a3.v[k] : a2.v[k];
    r.v[k] = b1 ? hi : lo;
  }
}

// p <- 2 p (dbl-2009-l); q = lane index within the quad
DEV void g1_dbl_quad(G1J& p, uint32_t q) {
  Fp a, b, r, A, B, YZ;
  fp_sel4(a, q, p.x, p.y, p.y, p.y);  // X^2 | Y^2 | Y Z | Y Z
  fp_sel4(b, q, p.x, p.y, p.z, p.z);
  fp_mul(r, a, b);
  fp_quad_bcast<0>(A, r);
  fp_quad_bcast<1>(B, r);
  fp_quad_bcast<2>(YZ, r);
  Fp E, S;
  fp_dbl(E, A);  // E = 3 A
  fp_add(E, E, A);
  fp_add(S, p.x, B);
  fp_sel4(a, q, E, B, S, S);  // F = E^2 | C = B^2 | T = (X + B)^2
  fp_sqr(r, a);
  Fp F, C;
  fp_quad_bcast<0>(F, r);
  fp_quad_bcast<1>(C, r);
  fp_quad_bcast<2>(S, r);
  fp_sub(S, S, A);  // D = 2 (T - A - C)
  fp_sub(S, S, C);
  fp_dbl(S, S);
  fp_sub(p.x, F, S);  // X3 = F - 2 D
  fp_sub(p.x, p.x, S);
  fp_sub(S, S, p.x);  // Y3 = E (D - X3) - 8 C
  fp_mul(p.y, E, S);
  fp_dbl(C, C);
  fp_dbl(C, C);
  fp_dbl(C, C);
  fp_sub(p.y, p.y, C);
  fp_dbl(p.z, YZ);  // Z3 = 2 Y Z (the identity keeps Z = 0)
}

// p <- p + o (add-2007-bl, both Jacobian, o not aliasing p)
DEV void g1_add_quad(G1J& p, const G1J& o, uint32_t q) {
  if (fp_is_zero(o.z)) return;
  if (fp_is_zero(p.z)) {
    p = o;
    return;
  }
  Fp a, b, r, Z1Z1, Z2Z2, U1, S1, H, R;
  fp_sel4(a, q, p.z, o.z, p.y, o.y);  // Z1^2 | Z2^2 | Y1 Z2 | Y2 Z1
  fp_sel4(b, q, p.z, o.z, o.z, p.z);
  fp_mul(r, a, b);
  fp_quad_bcast<0>(Z1Z1, r);
  fp_quad_bcast<1>(Z2Z2, r);
  fp_quad_bcast<2>(S1, r);
  fp_quad_bcast<3>(R, r);
  fp_sel4(a, q, p.x, o.x, S1, R);  // U1 | U2 | S1 | S2
  fp_sel4(b, q, Z2Z2, Z1Z1, Z2Z2, Z1Z1);
  fp_mul(r, a, b);
  fp_quad_bcast<0>(U1, r);
  fp_quad_bcast<1>(H, r);
  fp_quad_bcast<2>(S1, r);
  fp_quad_bcast<3>(R, r);
  fp_sub(H, H, U1);  // H = U2 - U1
  fp_sub(R, R, S1);  // S2 - S1
  if (fp_is_zero(H)) {
    if (fp_is_zero(R))
      g1_dbl_quad(p, q);
    else
      pt_set_inf(p);
    return;
  }
  fp_dbl(R, R);  // r = 2 (S2 - S1)
  fp_add(b, p.z, o.z);
  fp_dbl(a, H);
  fp_sel4(a, q, a, b, R, R);  // I = (2 H)^2 | (Z1 + Z2)^2 | r^2
  fp_sqr(r, a);
  Fp I, R2;
  fp_quad_bcast<0>(I, r);
  fp_quad_bcast<1>(b, r);
  fp_quad_bcast<2>(R2, r);
  fp_sub(b, b, Z1Z1);  // (Z1 + Z2)^2 - Z1Z1 - Z2Z2
  fp_sub(b, b, Z2Z2);
  fp_sel4(a, q, H, U1, b, b);  // J = H I | V = U1 I | Z3 = (..) H
  fp_sel4(b, q, I, I, H, H);
  fp_mul(r, a, b);
  Fp Jv, V;
  fp_quad_bcast<0>(Jv, r);
  fp_quad_bcast<1>(V, r);
  fp_quad_bcast<2>(p.z, r);
  fp_sub(p.x, R2, Jv);  // X3 = r^2 - J - 2 V
  fp_sub(p.x, p.x, V);
  fp_sub(p.x, p.x, V);
  fp_sub(V, V, p.x);
  fp_sel4(a, q, R, S1, R, R);  // r (V - X3) | S1 J
  fp_sel4(b, q, V, Jv, V, V);
  fp_mul(r, a, b);
  fp_quad_bcast<0>(V, r);
  fp_quad_bcast<1>(Jv, r);
  fp_dbl(Jv, Jv);  // Y3 = r (V - X3) - 2 S1 J
  fp_sub(p.y, V, Jv);
}

// ------------------------------------------------------------------ halving step
// Step s (1..K), M = B >> s: X_i += X_{i+M} for i in [0, M) and, for
// u = 1..s-1, in [B >> u, (B >> u) + M); reads and writes of one step are
// disjoint (the ranges written start at multiples of 2M, those read at an
// odd multiple of M).  A step holds one addition per element and few
// elements, so each addition runs on a quad (g1_add_quad).
__global__ __launch_bounds__(64) void k_pip_halve(PipShape sh, uint32_t step, uint32_t M,
                                                 uint32_t n, uint32_t* __restrict__ X) {
  pip_priority();
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t idx = t >> 2, q = t & 3;
  if (idx >= n) return;  // whole quads leave together
  const uint32_t per = step * M;
  const uint32_t jw = idx / per, rq = idx % per;
  const uint32_t u = rq / M, r = rq % M;
  const uint32_t i = (u ? (sh.B >> u) : 0u) + r;
  uint32_t* base = X + (size_t)jw * sh.B * G1J_WORDS;
  G1J a, b;
  load_g1j(a, base + (size_t)i * G1J_WORDS);
  load_g1j(b, base + (size_t)(i + M) * G1J_WORDS);
  g1_add_quad(a, b, q);
  if (q == 0) store_g1j(base + (size_t)i * G1J_WORDS, a);
}
void launch_pip_halve(hipStream_t s, const PipShape& sh, uint32_t step, const PipBuffers& bf) {
  if (!sh.terms || step < 1 || step > sh.K) return;
  const uint32_t M = sh.B >> step;
  const uint32_t n = sh.jobs * sh.W * step * M;
  hipLaunchKernelGGL(k_pip_halve, GRID(4 * (size_t)n, 64), 0, s, sh, step, M, n, bf.buckets);
}

// One block per MSM, one quad per window: quad w forms window w's sum
// X_0 + sum_j 2^j X_{2^j} (Horner over j), then quad 0 combines the windows,
// c doublings apart.
constexpr uint32_t PIP_FINAL_MAX_W = 33;  // c = 4
__global__ __launch_bounds__(256) void k_pip_final(PipShape sh, const uint32_t* __restrict__ job_out,
                                                  const uint32_t* __restrict__ X,
                                                  uint32_t* __restrict__ partials) {
  pip_priority();
  __shared__ uint32_t win[PIP_FINAL_MAX_W * G1J_WORDS];
  const uint32_t job = blockIdx.x, w = threadIdx.x >> 2, q = threadIdx.x & 3;
  if (w < sh.W) {
    const uint32_t* base = X + (size_t)(job * sh.W + w) * sh.B * G1J_WORDS;
    G1J acc, x;
    pt_set_inf(acc);
#pragma unroll 1
    for (int j = (int)sh.K - 1; j >= 0; j--) {
      g1_dbl_quad(acc, q);
      load_g1j(x, base + (size_t)(1u << j) * G1J_WORDS);
      g1_add_quad(acc, x, q);
    }
    load_g1j(x, base);
    g1_add_quad(acc, x, q);
    if (q == 0) store_g1j(win + w * G1J_WORDS, acc);
  }
  __syncthreads();
  if (w == 0) {
    G1J r, x;
    load_g1j(r, win + (sh.W - 1) * G1J_WORDS);
#pragma unroll 1
    for (int v = (int)sh.W - 2; v >= 0; v--) {
#pragma unroll 1
      for (uint32_t d = 0; d < sh.c; d++) g1_dbl_quad(r, q);
      load_g1j(x, win + v * G1J_WORDS);
      g1_add_quad(r, x, q);
    }
    if (q == 0) store_g1j(partials + (size_t)job_out[job] * G1J_WORDS, r);
  }
}
void launch_pip_final(hipStream_t s, const PipShape& sh, const uint32_t* job_out,
                      const PipBuffers& bf, uint32_t* partials) {
  if (!sh.terms) return;
  if (sh.W > PIP_FINAL_MAX_W) throw std::runtime_error("pippenger: more windows than the final stage holds");
  const uint32_t threads = (4 * sh.W + 63) / 64 * 64;
  hipLaunchKernelGGL(k_pip_final, dim3(sh.jobs), dim3(threads), 0, s, sh, job_out, bf.buckets, partials);
}

}  // namespace cubbs
