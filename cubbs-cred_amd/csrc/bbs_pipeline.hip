// Batched whole-operation pipelines (Sign, Verify, ProofGen, ProofVerify)
// built from generic stages: decode, term fill, variable-base GLV
// multiplication, reduction of partial sums, pairing checks, and the
// operation-specific hash/finalise steps.  Every kernel maps one lane to one
// job and is bounds-checked against its job count.
//
// Arrays shared by all stages of a batch:
//   scalars[] : canonical Fr (8 words); [0, T) are the fixed-base term scalars
//   points[]  : affine G1 (Montgomery, 2 x 14 words) with pinf[] identity flags
//   partials[]: Jacobian G1 partial sums
//   status[]  : per item, 0 ok, 1 invalid input / crypto failure
//   gate[]    : per item, extra condition of the pairing check (challenge)
#include "bbs_internal.h"
#include "dev_io.hpp"
#include "pairing12.hpp"

namespace cubbs {

// Every kernel of the pipelines launches single-wave (64-lane) workgroups: a
// pairing launch of another batch can hold most VGPRs of every SIMD, leaving
// one or two wave slots per CU, and a multi-wave workgroup (which needs all
// its waves on one CU) would wait until those pairing waves retire.
#define GRID(n, b) dim3(((n) + (b) - 1) / (b)), dim3(b)

__global__ void k_fill_status(uint32_t n, uint8_t* status, uint8_t* gate, uint8_t gv) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  status[i] = 0;
  gate[i] = gv;
}
void launch_fill_status(hipStream_t s, uint32_t n, uint8_t* status, uint8_t* gate, uint8_t gv) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_status, GRID(n, 64), 0, s, n, status, gate, gv);
}

// ------------------------------------------------------------------ transfers
// Copies between pinned host memory and device memory by a kernel on the
// batch's stream (the device reads and writes host memory directly), so the
// host thread only enqueues a launch: 16-byte accesses when both ends allow.
__global__ __launch_bounds__(256) void k_copy(uint8_t* __restrict__ dst,
                                             const uint8_t* __restrict__ src, size_t n) {
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i >= n) return;
  if (i + 16 <= n && ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0)) {
    *reinterpret_cast<uint4*>(dst + i) = *reinterpret_cast<const uint4*>(src + i);
  } else {
    const size_t e = i + 16 < n ? i + 16 : n;
    for (size_t k = i; k < e; k++) dst[k] = src[k];
  }
}
void launch_copy(hipStream_t s, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  const size_t lanes = (bytes + 15) / 16;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((lanes + 63) / 64)), dim3(64), 0, s,
                     static_cast<uint8_t*>(dst), static_cast<const uint8_t*>(src), bytes);
}

// ------------------------------------------------------------------ decoding
// check_subgroup = 0 leaves the prime-order subgroup check to the lanes that
// k_var_mul appends for these jobs.
__global__ __launch_bounds__(64, 2) void k_decode_g1(const DecodeG1Job* __restrict__ jobs, uint32_t n,
                                                 const uint8_t* __restrict__ blob,
                                                 uint32_t* __restrict__ points,
                                                 uint8_t* __restrict__ pinf,
                                                 uint8_t* __restrict__ status,
                                                 uint32_t check_subgroup) {
  wave_priority<PRIO_EARLY>();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DecodeG1Job j = jobs[i];
  G1J p;
  bool ok = g1_decompress(p, blob + j.off, false);
  G1A a;
  const bool finite = ok && !pt_is_inf(p);
  if (finite) {
    a.x = p.x;
    a.y = p.y;
    pinf[j.out] = 0;
  } else {
    fp_zero(a.x);
    fp_zero(a.y);
    pinf[j.out] = 1;
  }
  uint32_t* const at = points + (size_t)j.out * G1A_WORDS;
  store_g1a(at, a);
  // the check reads the point back from memory rather than keeping it live
  if (finite && check_subgroup) ok = g1_in_subgroup_at(at);
  if (!ok) status[j.item] = 1;
}
void launch_decode_g1(hipStream_t s, const DecodeG1Job* jobs, uint32_t n, const uint8_t* blob,
                      uint32_t* points, uint8_t* pinf, uint8_t* status, bool check_subgroup) {
  if (!n) return;
  hipLaunchKernelGGL(k_decode_g1, GRID(n, 64), 0, s, jobs, n, blob, points, pinf, status,
                     check_subgroup ? 1u : 0u);
}

__global__ __launch_bounds__(256) void k_decode_fr(const DecodeFrJob* __restrict__ jobs, uint32_t n,
                                                  const uint8_t* __restrict__ blob,
                                                  uint32_t* __restrict__ scalars,
                                                  uint8_t* __restrict__ status) {
  wave_priority<PRIO_EARLY>();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DecodeFrJob j = jobs[i];
  Fr s;
  fr_from_be_bytes_raw(s, blob + j.off);
  if (j.flags & 2) {
    // a 256-bit value is below 3r
    fr_sub_if_geq(s);
    fr_sub_if_geq(s);
  } else if (!fr_canonical(s)) {
    status[j.item] = 1;
    fr_zero(s);
  }
  if (j.flags & 4) {
    Fr m;
    fr_from_be_bytes_raw(m, blob + j.mul_off);
    fr_to_mont(m, m);
    fr_mul(s, s, m);  // canonical * Montgomery -> canonical product
  }
  if (j.flags & 1) fr_neg(s, s);
  store_fr(scalars + (size_t)j.out * FR_WORDS, s);
}

void launch_decode_fr(hipStream_t s, const DecodeFrJob* jobs, uint32_t n, const uint8_t* blob,
                      uint32_t* scalars, uint8_t* status) {
  if (!n) return;
  hipLaunchKernelGGL(k_decode_fr, GRID(n, 64), 0, s, jobs, n, blob, scalars, status);
}

// ------------------------------------------------------------------ term scalars
__global__ __launch_bounds__(256) void k_term_fill(const TermFill* __restrict__ jobs, uint32_t n,
                                                  uint32_t* __restrict__ scalars,
                                                  const uint32_t* __restrict__ sum_pairs) {
  wave_priority<PRIO_EARLY>();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  TermFill j = jobs[i];
  Fr r;
  if (j.mode == 2) {
    fr_zero(r);
    r.v[0] = 1;
  } else if (j.mode == 3) {
    fr_zero(r);
    for (uint32_t k = j.a; k < j.a + j.b; k++) {
      const uint32_t x = sum_pairs[2 * k], y = sum_pairs[2 * k + 1];
      Fr t;
      load_fr(t, scalars + (size_t)x * FR_WORDS);
      if (y != TERM_SUM_ONE) {
        Fr m;
        load_fr(m, scalars + (size_t)y * FR_WORDS);
        fr_to_mont(m, m);
        fr_mul(t, t, m);
      }
      fr_add(r, r, t);
    }
  } else {
    load_fr(r, scalars + (size_t)j.a * FR_WORDS);
    if (j.mode == 1) {
      Fr b;
      load_fr(b, scalars + (size_t)j.b * FR_WORDS);
      fr_to_mont(b, b);
      fr_mul(r, r, b);  // canonical * Montgomery -> canonical product
    }
  }
  store_fr(scalars + (size_t)j.dst * FR_WORDS, r);
}
void launch_term_fill(hipStream_t s, const TermFill* jobs, uint32_t n, uint32_t* scalars,
                      const uint32_t* sum_pairs) {
  if (!n) return;
  hipLaunchKernelGGL(k_term_fill, GRID(n, 64), 0, s, jobs, n, scalars, sum_pairs);
}

// ------------------------------------------------------------------ variable-base products
// Signed 4-bit windows of a 128-bit GLV half e: E' = e + sum_{i<=32} 8 * 16^i
// (five words), digit i = nibble i of E' minus 8, in [-8, 7]; digit 32 is
// the carry (0 or 1).  sum_i d_i 16^i = e.
DEV void var_digits(uint32_t ed[5], const uint32_t e[4]) {
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    c += (uint64_t)(i < 4 ? e[i] : 0u) + (i < 4 ? 0x88888888u : 0x8u);
    ed[i] = (uint32_t)c;
    c >>= 32;
  }
}
DEV int var_digit(const uint32_t ed[5], int i) {
  const int w = i >> 3;
  uint32_t v = ed[0];
#pragma unroll
  for (int q = 1; q < 5; q++) v = w == q ? ed[q] : v;
  return (int)((v >> ((i & 7) * 4)) & 15u) - 8;
}

// tab[j] = (j + 1) q, j = 0..7, affine: 2q..8q in Jacobian form (jw, 7 points)
// then normalised with one inversion (Montgomery's trick; the running
// products are parked in tab[1..7] until the backward pass overwrites them).
DEV void var_table(const G1A& q, uint32_t* tab, uint32_t* jw) {
  store_g1a(tab, q);
  G1J a, b;
  pt_from_affine(a, q);
  g1_dbl_in(a);                          // 2q
  store_g1j(jw + 0 * G1J_WORDS, a);
  b = a;
  g1_dbl_in(b);                          // 4q
  store_g1j(jw + 2 * G1J_WORDS, b);
  g1_madd_in(a, q);                      // 3q
  store_g1j(jw + 1 * G1J_WORDS, a);
  g1_dbl_in(a);                          // 6q
  store_g1j(jw + 4 * G1J_WORDS, a);
  g1_madd_in(a, q);                      // 7q
  store_g1j(jw + 5 * G1J_WORDS, a);
  a = b;
  g1_madd_in(a, q);                      // 5q
  store_g1j(jw + 3 * G1J_WORDS, a);
  g1_dbl_in(b);                          // 8q
  store_g1j(jw + 6 * G1J_WORDS, b);
  Fp run, z;
#pragma unroll 1
  for (int j = 0; j < 7; j++) {
    ld_fp(z, jw + j * G1J_WORDS + 2 * FPW);
    if (j)
      fp_mul(run, run, z);
    else
      run = z;
    st_fp(tab + (j + 1) * G1A_WORDS, run);
  }
  Fp inv;
  fp_inv(inv, run);
#pragma unroll 1
  for (int j = 6; j >= 0; j--) {
    Fp zi, pre, zi2, zi3, X, Y;
    ld_fp(z, jw + j * G1J_WORDS + 2 * FPW);
    if (j) {
      ld_fp(pre, tab + j * G1A_WORDS);
      fp_mul(zi, inv, pre);
    } else {
      zi = inv;
    }
    fp_mul(inv, inv, z);
    fp_sqr(zi2, zi);
    fp_mul(zi3, zi2, zi);
    ld_fp(X, jw + j * G1J_WORDS);
    ld_fp(Y, jw + j * G1J_WORDS + FPW);
    G1A o;
    fp_mul(o.x, X, zi2);
    fp_mul(o.y, Y, zi3);
    store_g1a(tab + (j + 1) * G1A_WORDS, o);
  }
}

DEV void var_entry(G1A& e, const uint32_t* tab, int d) {
  const int a = d < 0 ? -d : d;
  load_g1a(e, tab + (a - 1) * G1A_WORDS);
}

// [k] P for an affine P (not the identity) on one lane: k = k1 + k2 lambda,
// k1 P + k2 phi(P) evaluated from the top window down with 4 shared
// doublings and at most two mixed additions per window (uniform control
// flow across the lanes of a wave, unlike a bit-by-bit ladder).  The entries
// of the next window are loaded before the doublings.
DEV void var_mul_one(G1J& r, const G1A& p, const uint32_t k[8], uint32_t* ws) {
  uint32_t k1[4], k2[4];
  glv_split(k1, k2, k);
  uint32_t* t1 = ws;
  uint32_t* t2 = ws + VAR_TABLE_WORDS;
  var_table(p, t1, ws + 2 * VAR_TABLE_WORDS);
  {
    Fp beta;
    fp_load_const(beta, bls::GLV_BETA);
#pragma unroll 1
    for (int j = 0; j < 8; j++) {  // phi(j P) = (beta x, y)
      G1A e;
      load_g1a(e, t1 + j * G1A_WORDS);
      fp_mul(e.x, e.x, beta);
      store_g1a(t2 + j * G1A_WORDS, e);
    }
  }
  uint32_t d1[5], d2[5];
  var_digits(d1, k1);
  var_digits(d2, k2);
  int top = 32;
  while (top >= 0 && var_digit(d1, top) == 0 && var_digit(d2, top) == 0) top--;
  G1J acc;
  pt_set_inf(acc);
  // one copy each of the doubling and of the addition in the loop body keeps
  // it within the instruction cache
#pragma unroll 1
  for (int i = top; i >= 0; i--) {
    const int a = var_digit(d1, i), b = var_digit(d2, i);
    G1A ea, eb;
    if (a) var_entry(ea, t1, a);
    if (b) var_entry(eb, t2, b);
    if (i != top) {
#pragma unroll 1
      for (int s = 0; s < 4; s++) g1_dbl_in(acc);
    }
#pragma unroll 1
    for (int s = 0; s < 2; s++) {
      const int d = s ? b : a;
      if (d) {
        G1A e;
        e.x = s ? eb.x : ea.x;
        e.y = s ? eb.y : ea.y;
        if (d < 0) fp_neg(e.y, e.y);
        g1_madd_in(acc, e);
      }
    }
  }
  r = acc;
}

// The tables of a multi-term job in its workspace ws: 1..8 P_t and 1..8
// phi(P_t) of every term t that is not the identity (the identity's digits
// are all zero, so its tables are never read).  Job fields are read through
// the pointer (a copy indexed by the loop counter would sit in scratch).
DEV void var_sum_tables(const VarSum* __restrict__ jp, const uint32_t* __restrict__ points,
                        const uint8_t* __restrict__ pinf, uint32_t* ws) {
  const uint32_t count = jp->count;
  uint32_t* const jw = ws + count * 2 * VAR_TABLE_WORDS;
#pragma unroll 1
  for (uint32_t t = 0; t < count; t++) {
    const uint32_t pt = jp->point[t];
    if (pinf[pt]) continue;
    G1A p;
    load_g1a(p, points + (size_t)pt * G1A_WORDS);
    uint32_t* const t1 = ws + t * 2 * VAR_TABLE_WORDS;
    uint32_t* const t2 = t1 + VAR_TABLE_WORDS;
    var_table(p, t1, jw);
    Fp beta;
    fp_load_const(beta, bls::GLV_BETA);
#pragma unroll 1
    for (int e = 0; e < 8; e++) {  // phi(e P) = (beta x, y)
      G1A q;
      load_g1a(q, t1 + e * G1A_WORDS);
      fp_mul(q.x, q.x, beta);
      store_g1a(t2 + e * G1A_WORDS, q);
    }
  }
}

// [sum_t k_t P_t] on one lane, k_t = scalars[sc[t]], over the tables that
// var_sum_tables left in ws: the signed 4-bit digits of both GLV halves of
// every term in LDS (word w of term t at dg[(10 t + w) * 64 + lane]), then
// one chain of doublings from the highest non-zero window down, with at
// most 2 * count mixed additions per window.
DEV void var_sum_chain(G1J& r, const VarSum* __restrict__ jp, const uint32_t* __restrict__ sc,
                       const uint8_t* __restrict__ pinf, const uint32_t* __restrict__ scalars,
                       const uint32_t* ws, uint32_t* dg, uint32_t lane) {
  const uint32_t count = jp->count;
  int top = -1;
#pragma unroll 1
  for (uint32_t t = 0; t < count; t++) {
    uint32_t d1[5] = {0x88888888u, 0x88888888u, 0x88888888u, 0x88888888u, 8u};
    uint32_t d2[5] = {0x88888888u, 0x88888888u, 0x88888888u, 0x88888888u, 8u};  // all digits 0
    if (!pinf[jp->point[t]]) {
      Fr k;
      load_fr(k, scalars + (size_t)sc[t] * FR_WORDS);
      uint32_t k1[4], k2[4];
      glv_split(k1, k2, k.v);
      var_digits(d1, k1);
      var_digits(d2, k2);
    }
#pragma unroll
    for (int w = 0; w < 5; w++) {
      dg[(10 * t + w) * 64 + lane] = d1[w];
      dg[(10 * t + 5 + w) * 64 + lane] = d2[w];
    }
    int tt = 32;
    while (tt >= 0 && var_digit(d1, tt) == 0 && var_digit(d2, tt) == 0) tt--;
    top = tt > top ? tt : top;
  }
  G1J acc;
  pt_set_inf(acc);
#pragma unroll 1
  for (int i = top; i >= 0; i--) {
    if (i != top) {
#pragma unroll 1
      for (int s = 0; s < 4; s++) g1_dbl_in(acc);
    }
#pragma unroll 1
    for (uint32_t h = 0; h < 2 * count; h++) {  // (term, GLV half)
      const uint32_t w = dg[(5 * h + (uint32_t)(i >> 3)) * 64 + lane];
      const int d = (int)((w >> ((i & 7) * 4)) & 15u) - 8;
      if (d) {
        G1A e;
        var_entry(e, ws + h * VAR_TABLE_WORDS, d);
        if (d < 0) fp_neg(e.y, e.y);
        g1_madd_in(acc, e);
      }
    }
  }
  r = acc;
}

// Lanes [0, ns): the multi-term jobs (ns = n_sums rounded up to whole waves,
// so no wave mixes them with single products); a dual job forms its second
// sum over the same tables, a gated job returns at once when valid[gate] is
// set, and with sum_list only the jobs sum_list[0 .. *sum_count) run.  Lanes
// [ns, ns + n): one product per job (workspace ws + t * VAR_WS_WORDS).
// Lanes after those, up to n_sub more: the prime-order subgroup check of the
// points decoded by jobs sub[] without it, run beside the products; a
// failure marks the item's status (the products of such a point are
// computed but never used).
__global__ __launch_bounds__(64, 2) void k_var_mul(const VarJob* __restrict__ jobs, uint32_t n,
                                               const uint32_t* __restrict__ points,
                                               const uint8_t* __restrict__ pinf,
                                               const uint32_t* __restrict__ scalars,
                                               uint32_t* __restrict__ partials,
                                               uint32_t* __restrict__ ws,
                                               const DecodeG1Job* __restrict__ sub, uint32_t n_sub,
                                               uint8_t* __restrict__ status,
                                               const VarSum* __restrict__ sums, uint32_t n_sums,
                                               uint32_t* __restrict__ sum_ws,
                                               const uint8_t* __restrict__ valid,
                                               const uint32_t* __restrict__ sum_list,
                                               const uint32_t* __restrict__ sum_count) {
  wave_priority<PRIO_EARLY>();
  extern __shared__ uint32_t dg[];  // 10 * VAR_TERMS_MAX * 64 words when n_sums > 0
  const uint32_t ns = (n_sums + 63) & ~63u;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ns) {
    if (g >= (sum_list ? *sum_count : n_sums)) return;
    const VarSum* const jp = sums + (sum_list ? sum_list[g] : g);
    const uint32_t flags = jp->flags;
    if (!sum_list && (flags & VAR_SUM_GATED) && valid[jp->gate]) return;
    uint32_t* const w = sum_ws + jp->ws;
    var_sum_tables(jp, points, pinf, w);
    G1J r;
    var_sum_chain(r, jp, jp->scalar, pinf, scalars, w, dg, threadIdx.x);
    store_g1j(partials + (size_t)jp->partial * G1J_WORDS, r);
    if (flags & VAR_SUM_DUAL) {
      var_sum_chain(r, jp, jp->scalar2, pinf, scalars, w, dg, threadIdx.x);
      store_g1j(partials + (size_t)jp->partial2 * G1J_WORDS, r);
    }
    return;
  }
  const uint32_t t = g - ns;
  if (t >= n) {
    const uint32_t k = t - n;
    if (k >= n_sub) return;
    const DecodeG1Job j = sub[k];
    if (pinf[j.out]) return;  // the identity, or not decoded (status already set)
    if (!g1_in_subgroup_at(points + (size_t)j.out * G1A_WORDS)) status[j.item] = 1;
    return;
  }
  const VarJob j = jobs[t];
  G1J r;
  pt_set_inf(r);
  if (!pinf[j.point]) {
    G1A p;
    load_g1a(p, points + (size_t)j.point * G1A_WORDS);
    Fr k;
    load_fr(k, scalars + (size_t)j.scalar * FR_WORDS);
    var_mul_one(r, p, k.v, ws + (size_t)t * VAR_WS_WORDS);
  }
  store_g1j(partials + (size_t)j.partial * G1J_WORDS, r);
}
void launch_var_mul(hipStream_t s, const VarJob* jobs, uint32_t n, const uint32_t* points,
                    const uint8_t* pinf, const uint32_t* scalars, uint32_t* partials, uint32_t* ws,
                    const DecodeG1Job* sub, uint32_t n_sub, uint8_t* status, const VarSum* sums,
                    uint32_t n_sums, uint32_t* sum_ws, const uint8_t* valid, const uint32_t* sum_list,
                    const uint32_t* sum_count) {
  if (!n && !n_sub && !n_sums) return;
  const uint32_t ns = (n_sums + 63) & ~63u;
  const size_t lds = n_sums ? 4ull * 10 * VAR_TERMS_MAX * 64 : 0;
  hipLaunchKernelGGL(k_var_mul, GRID(ns + n + n_sub, 64), lds, s, jobs, n, points, pinf, scalars,
                     partials, ws, sub, n_sub, status, sums, n_sums, sum_ws, valid, sum_list, sum_count);
}

// ------------------------------------------------------------------ reduction
// points[o] = acc in affine coordinates (pinf[o] marks the identity).
DEV void store_affine_sum(const G1J& acc, uint32_t o, uint32_t* points, uint8_t* pinf) {
  G1A a;
  if (pt_is_inf(acc)) {
    pinf[o] = 1;
    fp_zero(a.x);
    fp_zero(a.y);
  } else {
    pinf[o] = 0;
    G1J t = acc;  // the out-of-line conversion takes a copy's address, not acc's
    pt_to_affine(a, t);
  }
  store_g1a(points + (size_t)o * G1A_WORDS, a);
}

// The sum of partials [begin, begin + count) on the RL consecutive lanes
// q = 0 .. RL - 1 (q = threadIdx.x % RL) of one 64-lane block's sum: lane q
// adds every RL-th partial from begin + q, then log2(RL) halving steps
// through LDS (xs: 64 points) leave the sum in lane q = 0.  Every lane of the
// block calls it (it synchronises the block); lanes with live == false add
// nothing.  A sum of count partials takes ceil(count / RL) + log2(RL)
// additions one after the other instead of count.
template <uint32_t RL>
DEV void coop_sum(G1J& acc, bool live, uint32_t begin, uint32_t count,
                  const uint32_t* __restrict__ partials, uint32_t* xs) {
  const uint32_t q = threadIdx.x % RL;
  pt_set_inf(acc);
  if (live) {
    for (uint32_t k = q; k < count; k += RL) {
      G1J p;
      load_g1j(p, partials + (size_t)(begin + k) * G1J_WORDS);
      g1_add_in(acc, p);
    }
  }
  for (uint32_t h = RL / 2; h > 0; h >>= 1) {
    store_g1j(xs + threadIdx.x * G1J_WORDS, acc);
    __syncthreads();
    if (live && q < h) {
      G1J p;
      load_g1j(p, xs + (threadIdx.x + h) * G1J_WORDS);  // lane q + h of the same sum
      g1_add_in(acc, p);
    }
    __syncthreads();
  }
}

// One lane per group (at most Plan::FINAL_MAX partials): the launches are
// wide, and the affine conversion at the end costs a wave per 64 sums.
__global__ __launch_bounds__(64) void k_reduce(const ReduceGroup* __restrict__ groups,
                                              const uint32_t* __restrict__ out_point, uint32_t n,
                                              const uint32_t* __restrict__ partials,
                                              uint32_t* __restrict__ points,
                                              uint8_t* __restrict__ pinf) {
  wave_priority<PRIO_MID>();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ReduceGroup g = groups[i];
  G1J acc;
  pt_set_inf(acc);
  for (uint32_t k = 0; k < g.count; k++) {
    G1J p;
    load_g1j(p, partials + (size_t)(g.begin + k) * G1J_WORDS);
    g1_add_in(acc, p);
  }
  store_affine_sum(acc, out_point[i], points, pinf);
}
void launch_reduce(hipStream_t s, const ReduceGroup* groups, const uint32_t* out_point, uint32_t n,
                   const uint32_t* partials, uint32_t* points, uint8_t* pinf) {
  if (!n) return;
  hipLaunchKernelGGL(k_reduce, GRID(n, 64), 0, s, groups, out_point, n, partials, points, pinf);
}

// The sums of the per-item checks of aggregated groups: only for the items of
// groups whose combined check failed (valid[gate] == 0, written by the group
// pairing launch), or the jobs list[0 .. *count) selected as such; one sum on
// GATED_LANES lanes (up to (L + 2) / 4 comb partials for an item of L
// messages: a long serial chain on one lane, in a launch of few sums).
constexpr uint32_t GATED_LANES = 4;
__global__ __launch_bounds__(64) void k_reduce_gated(const GatedReduce* __restrict__ jobs, uint32_t n,
                                                    const uint32_t* __restrict__ partials,
                                                    uint32_t* __restrict__ points,
                                                    uint8_t* __restrict__ pinf,
                                                    const uint8_t* __restrict__ valid,
                                                    const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ count) {
  __shared__ uint32_t xs[64 * G1J_WORDS];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t i = t / GATED_LANES;
  bool live = i < (list ? *count : n);
  GatedReduce g{0, 0, REDUCE_NO_EXTRA, 0, 0};
  if (live) {
    g = jobs[list ? list[i] : i];
    if (!list && valid[g.gate]) live = false;  // the same for all lanes of the sum
  }
  G1J acc;
  coop_sum<GATED_LANES>(acc, live, g.begin, g.count, partials, xs);
  if (!live || t % GATED_LANES != 0) return;
  if (g.extra != REDUCE_NO_EXTRA) {
    G1J p;
    load_g1j(p, partials + (size_t)g.extra * G1J_WORDS);
    g1_add_in(acc, p);
  }
  store_affine_sum(acc, g.out, points, pinf);
}
void launch_reduce_gated(hipStream_t s, const GatedReduce* jobs, uint32_t n, const uint32_t* partials,
                         uint32_t* points, uint8_t* pinf, const uint8_t* valid, const uint32_t* list,
                         const uint32_t* count) {
  if (!n) return;
  hipLaunchKernelGGL(k_reduce_gated, GRID(n * GATED_LANES, 64), 0, s, jobs, n, partials, points, pinf,
                     valid, list, count);
}

// Lanes over the concatenated jobs of the three kinds; one atomic per wave
// and kind reserves the places of its live jobs.
__global__ __launch_bounds__(64) void k_gate_select(GateLists g, const uint8_t* __restrict__ valid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t k = 3, j = i;
  if (j < g.n[0]) {
    k = 0;
  } else if ((j -= g.n[0]) < g.n[1]) {
    k = 1;
  } else if ((j -= g.n[1]) < g.n[2]) {
    k = 2;
  }
  const bool need = k < 3 && valid[g.gate[k][j]] == 0;
  const uint32_t lane = threadIdx.x & 63;
#pragma unroll
  for (uint32_t q = 0; q < 3; q++) {
    const uint64_t m = __ballot(need && k == q);
    if (!m) continue;
    const uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(g.count + q, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, (int)first);
    if (need && k == q) g.list[q][base + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = j;
  }
}
void launch_gate_select(hipStream_t s, const GateLists& g, const uint8_t* valid) {
  const uint32_t n = g.n[0] + g.n[1] + g.n[2];
  if (!n) return;
  hipLaunchKernelGGL(k_gate_select, GRID(n, 64), 0, s, g, valid);
}

// One level of a partial-sum tree: each lane adds a short run of Jacobian
// partials into a new partial slot (large MSMs reduce in parallel levels).
__global__ __launch_bounds__(64) void k_reduce_partials(const ReduceGroup* __restrict__ groups,
                                                       const uint32_t* __restrict__ out, uint32_t n,
                                                       uint32_t* __restrict__ partials) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ReduceGroup g = groups[i];
  G1J acc;
  pt_set_inf(acc);
  for (uint32_t k = 0; k < g.count; k++) {
    G1J p;
    load_g1j(p, partials + (size_t)(g.begin + k) * G1J_WORDS);
    g1_add_in(acc, p);
  }
  store_g1j(partials + (size_t)out[i] * G1J_WORDS, acc);
}
void launch_reduce_partials(hipStream_t s, const ReduceGroup* groups, const uint32_t* out,
                            uint32_t n, uint32_t* partials) {
  if (!n) return;
  hipLaunchKernelGGL(k_reduce_partials, GRID(n, 64), 0, s, groups, out, n, partials);
}

__global__ __launch_bounds__(64) void k_compress(const CompressJob* __restrict__ jobs, uint32_t n,
                                                const uint32_t* __restrict__ points,
                                                const uint8_t* __restrict__ pinf,
                                                uint8_t* __restrict__ out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CompressJob j = jobs[i];
  uint8_t c[48];
  if (pinf[j.point]) {
    c[0] = 0xc0;
    for (int k = 1; k < 48; k++) c[k] = 0;
  } else {
    G1A a;
    load_g1a(a, points + (size_t)j.point * G1A_WORDS);
    g1_compress_affine(c, a);
  }
  for (int k = 0; k < 48; k++) out[j.off + k] = c[k];
}
void launch_compress(hipStream_t s, const CompressJob* jobs, uint32_t n, const uint32_t* points,
                     const uint8_t* pinf, uint8_t* out) {
  if (!n) return;
  hipLaunchKernelGGL(k_compress, GRID(n, 64), 0, s, jobs, n, points, pinf, out);
}

// ------------------------------------------------------------------ pairing check
// e(a, PK) * e(b, -G2) == 1 (identity arguments contribute 1).
//
// k_pair_lines prepares the item-dependent part of every Miller-loop line;
// the pairing kernel then runs each check on 12 lanes (pairing12.hpp), 5
// checks per 64-thread block.
//
// Line blocks of the pairing checks: one lane per (check, line, pairing)
// forms m = -lambda x_P and writes the line's 9 second factors
// (p12::line_values); a pairing with an identity point or an identity key
// (skipped by the check) gets the constant line 1.  With `group_valid`, the
// checks of items whose group check held (an earlier launch) are skipped:
// the pairing kernel accepts them without reading their blocks.
__global__ __launch_bounds__(256) void k_pair_lines(const PairJob* __restrict__ jobs, uint32_t n,
                                                   const uint32_t* __restrict__ points,
                                                   const uint8_t* __restrict__ pinf,
                                                   const uint8_t* __restrict__ pk_status,
                                                   const uint32_t* __restrict__ pk_lines,
                                                   const uint32_t* __restrict__ neg_g2_lines,
                                                   uint32_t* __restrict__ l2buf,
                                                   const uint8_t* __restrict__ group_valid,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * (uint32_t)(MILLER_LINES * 2)) return;
  const uint32_t slot = t / (MILLER_LINES * 2);
  if (list && slot >= *count) return;
  const uint32_t job = list ? list[slot] : slot;
  const uint32_t r = t % (MILLER_LINES * 2);
  const uint32_t idx = r >> 1, q = r & 1;
  const PairJob j = jobs[job];
  if (group_valid && (j.flags & PAIR_GROUPED) && group_valid[j.group]) return;
  const uint32_t* qlines = q ? (j.pk_b == PAIR_NEG_G2 ? neg_g2_lines
                                                       : pk_lines + (size_t)j.pk_b * PK_LINES_WORDS)
                             : pk_lines + (size_t)j.pk_slot * PK_LINES_WORDS;
  const uint32_t* line = qlines + (size_t)idx * MILLER_LINE_WORDS;
  Fp l0, l1, c0, c1, x, m0, m1;
  fp_load_const(l0, line);
  fp_load_const(l1, line + FPW);
  fp_load_const(c0, line + 2 * FPW);
  fp_load_const(c1, line + 3 * FPW);
  fp_load_const(x, points + (size_t)(q ? j.b : j.a) * G1A_WORDS);
  fp_mul(m0, l0, x);
  fp_neg(m0, m0);
  fp_mul(m1, l1, x);
  fp_neg(m1, m1);
  const uint8_t pks = pk_status[j.pk_slot];
  const uint8_t pkb = j.pk_b == PAIR_NEG_G2 ? 1 : pk_status[j.pk_b];
  const bool skip = q ? (pinf[j.b] || pkb == 2) : (pinf[j.a] || pks == 2);
  p12::line_values(l2buf + (size_t)slot * p12::L2_WORDS + ((size_t)idx * 2 + q) * p12::LV_N * FPW,
                   m0, m1, c0, c1, skip);
}
void launch_pair_lines(hipStream_t s, const PairJob* jobs, uint32_t n, const uint32_t* points,
                       const uint8_t* pinf, const uint8_t* pk_status, const uint32_t* pk_lines,
                       const uint32_t* neg_g2_lines, uint32_t* l2buf, const uint8_t* group_valid,
                       const uint32_t* list, const uint32_t* count) {
  if (!n) return;
  const uint32_t nl = n * (uint32_t)(MILLER_LINES * 2);
  hipLaunchKernelGGL(k_pair_lines, GRID(nl, 64), 0, s, jobs, n, points, pinf, pk_status, pk_lines,
                     neg_g2_lines, l2buf, group_valid, list, count);
}

// Per-item checks after group checks: a job accepted through its group (the
// item's own equation then holds except with probability 2^-128) or not live
// (an item already rejected, an invalid key, a failed challenge gate) gets
// its verdict here; the jobs that need their own pairing are listed
// contiguously (one atomic per wave, lanes in order), so the pairing launch
// holds only blocks of checks that run.  Left among many empty blocks, the
// few live ones were placed unevenly by the dispatcher (two on one SIMD),
// which lengthened the launch by about half a check.
__global__ __launch_bounds__(64) void k_pair_select(const PairJob* __restrict__ jobs, uint32_t n,
                                                   const uint8_t* __restrict__ pk_status,
                                                   uint8_t* __restrict__ status,
                                                   const uint8_t* __restrict__ gate,
                                                   uint8_t* __restrict__ valid,
                                                   uint32_t* __restrict__ list,
                                                   uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool need = false;
  if (i < n) {
    const PairJob j = jobs[i];
    const uint8_t pks = pk_status[j.pk_slot];
    const uint8_t pkb = j.pk_b == PAIR_NEG_G2 ? 1 : pk_status[j.pk_b];
    const bool live = status[j.item] == 0 && pks != 0 && pkb != 0 && gate[j.item];
    const bool accepted = live && (j.flags & PAIR_GROUPED) && valid[j.group];
    need = live && !accepted;
    if (!need) {
      if (pks == 0 || pkb == 0) status[j.item] = 1;
      valid[j.item] = accepted ? 1 : 0;
    }
  }
  const uint64_t m = __ballot(need);
  if (!m) return;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t first = (uint32_t)__ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == first) base = atomicAdd(count, (uint32_t)__popcll(m));
  base = (uint32_t)__shfl((int)base, (int)first);
  if (need) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1))] = i;
}
void launch_pair_select(hipStream_t s, const PairJob* jobs, uint32_t n, const uint8_t* pk_status,
                        uint8_t* status, const uint8_t* gate, uint8_t* valid, uint32_t* list,
                        uint32_t* count) {
  if (!n) return;
  hipLaunchKernelGGL(k_pair_select, GRID(n, 64), 0, s, jobs, n, pk_status, status, gate, valid, list,
                     count);
}

__global__ __launch_bounds__(64, 2) void k_pairing(const PairJob* __restrict__ jobs, uint32_t n,
                                               const uint32_t* __restrict__ points,
                                               const uint8_t* __restrict__ pinf,
                                               const uint8_t* __restrict__ pk_status,
                                               const uint32_t* __restrict__ l2buf,
                                               uint8_t* __restrict__ status,
                                               const uint8_t* __restrict__ gate,
                                               uint8_t* __restrict__ valid,
                                               const uint32_t* __restrict__ list,
                                               const uint32_t* __restrict__ count) {
  __shared__ uint32_t lds_all[p12::BLOCK_LDS_WORDS];
  const int lane = threadIdx.x;
  const int grp = lane / 12;
  const p12::Lane L = p12::lane_setup(lds_all, lane);
  // slot: this check's place in the launch (its line blocks); job: its record
  const uint32_t slot = blockIdx.x * p12::CHECKS_PER_BLOCK + grp;
  const bool mine = grp < p12::CHECKS_PER_BLOCK && slot < (list ? *count : n);
  const uint32_t job = list && mine ? list[slot] : slot;
  bool live = mine, accepted = false;
  PairJob j{};
  uint8_t pks = 1, pkb = 1;
  if (mine) {
    j = jobs[job];
    pks = pk_status[j.pk_slot];
    pkb = j.pk_b == PAIR_NEG_G2 ? 1 : pk_status[j.pk_b];
    live = status[j.item] == 0 && pks != 0 && pkb != 0 && gate[j.item];
    // the group's combined check (an earlier launch) passed: the item's own
    // equation holds except with probability 2^-128
    if (live && (j.flags & PAIR_GROUPED) && valid[j.group]) {
      accepted = true;
      live = false;
    }
  }
  // blocks with no live check leave together (uniform exit, no barrier skipped)
  if (!__syncthreads_or(live)) {
    if (mine && lane % 12 == 0) {
      if (pks == 0 || pkb == 0) status[j.item] = 1;
      valid[j.item] = accepted ? 1 : 0;
    }
    return;
  }
  // a skipped pairing (identity point or key) has the line 1 in its blocks
  // (k_pair_lines) and y = 0
  Fp yp0, yp1;
  fp_zero(yp0);
  fp_zero(yp1);
  const uint32_t* l2 = l2buf;
  if (live) {
    if (!(pinf[j.a] || pks == 2)) fp_load_const(yp0, points + (size_t)j.a * G1A_WORDS + FPW);
    if (!(pinf[j.b] || pkb == 2)) {
      fp_load_const(yp1, points + (size_t)j.b * G1A_WORDS + FPW);
      if (j.flags & PAIR_NEG_B) fp_neg(yp1, yp1);
    }
    l2 = l2buf + (size_t)slot * p12::L2_WORDS;
  }
  Fp f, r;
  p12::miller(f, l2, yp0, yp1, L);
  p12::final_exp(r, f, L);
  const bool one = p12::is_one(r, L);
  if (mine && lane % 12 == 0) {
    if (pks == 0 || pkb == 0) status[j.item] = 1;
    valid[j.item] = (accepted || (live && one)) ? 1 : 0;
  }
}
void launch_pairing(hipStream_t s, const PairJob* jobs, uint32_t n, const uint32_t* points,
                    const uint8_t* pinf, const uint8_t* pk_status, uint32_t* l2buf, uint8_t* status,
                    const uint8_t* gate, uint8_t* valid, const uint32_t* list, const uint32_t* count) {
  if (!n) return;
  hipLaunchKernelGGL(k_pairing, dim3((n + p12::CHECKS_PER_BLOCK - 1) / p12::CHECKS_PER_BLOCK),
                     dim3(64), 0, s, jobs, n, points, pinf, pk_status, l2buf, status, gate, valid, list,
                     count);
}

// ------------------------------------------------------------------ Sign
// core_sign (crates/cubbs-bbs-protocol/src/bbsplus/signature.rs:544-580):
// e = hash_to_scalar(sk || m_1 .. m_L || domain, api_id || "H2S_"); the term
// scalars [1, domain, m_1 .. m_L] are rescaled by 1 / (sk + e) so that the
// comb MSM yields A = B / (sk + e) directly (signature.rs:642-674).
__global__ __launch_bounds__(64) void k_sign_e(const SignJob* __restrict__ jobs, uint32_t n,
                                              const uint8_t* __restrict__ blob,
                                              const uint8_t* __restrict__ pk_status,
                                              uint32_t* __restrict__ scalars,
                                              uint8_t* __restrict__ status) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SignJob j = jobs[i];
  Fr sk;
  fr_from_be_bytes_raw(sk, blob + j.sk_off);
  if (!fr_canonical(sk) || pk_status[j.pk_slot] == 0) {
    status[j.item] = 1;
    return;
  }
  Sha256 s;
  sha_init_after_zpad(s);
  sha_scalar(s, sk);
  for (uint32_t k = 0; k < j.L; k++) {
    Fr m;
    load_fr(m, scalars + (size_t)(j.term + 2 + k) * FR_WORDS);
    sha_scalar(s, m);
  }
  Fr dom;
  load_fr(dom, scalars + (size_t)(j.term + 1) * FR_WORDS);
  sha_scalar(s, dom);
  xmd_b0_suffix(s, 48, DST_H2S, DST_H2S_LEN);
  uint32_t okm[12];
  xmd48_from_b0(okm, s.h, DST_H2S, DST_H2S_LEN);
  Fr e;
  scalar_from_okm48(e, okm);
  store_fr(scalars + (size_t)j.e_index * FR_WORDS, e);
  Fr se;
  fr_add(se, sk, e);
  if (fr_is_zero(se)) {
    status[j.item] = 1;
    return;
  }
  Fr inv;
  fr_to_mont(se, se);
  fr_inv(inv, se);  // Montgomery
  Fr t;
  fr_from_mont(t, inv);
  store_fr(scalars + (size_t)j.term * FR_WORDS, t);  // P1 term: 1 / (sk + e)
  for (uint32_t k = 1; k < j.L + 2; k++) {
    Fr x;
    load_fr(x, scalars + (size_t)(j.term + k) * FR_WORDS);
    fr_mul(x, x, inv);
    store_fr(scalars + (size_t)(j.term + k) * FR_WORDS, x);
  }
}
void launch_sign_e(hipStream_t s, const SignJob* jobs, uint32_t n, const uint8_t* blob,
                   const uint8_t* pk_status, uint32_t* scalars, uint8_t* status) {
  if (!n) return;
  hipLaunchKernelGGL(k_sign_e, GRID(n, 64), 0, s, jobs, n, blob, pk_status, scalars, status);
}

__global__ __launch_bounds__(64) void k_sign_finalize(const SignJob* __restrict__ jobs, uint32_t n,
                                                     const uint32_t* __restrict__ points,
                                                     const uint8_t* __restrict__ pinf,
                                                     const uint32_t* __restrict__ scalars,
                                                     uint8_t* __restrict__ status,
                                                     uint8_t* __restrict__ sig_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  SignJob j = jobs[i];
  if (status[j.item]) return;
  if (pinf[j.a_point]) {  // A == identity is rejected (signature.rs:664-668)
    status[j.item] = 1;
    return;
  }
  G1A a;
  load_g1a(a, points + (size_t)j.a_point * G1A_WORDS);
  uint8_t c[48];
  g1_compress_affine(c, a);
  uint8_t* o = sig_out + (size_t)80 * j.item;
  for (int k = 0; k < 48; k++) o[k] = c[k];
  Fr e;
  load_fr(e, scalars + (size_t)j.e_index * FR_WORDS);
  fr_to_be_bytes(o + 48, e);
}
void launch_sign_finalize(hipStream_t s, const SignJob* jobs, uint32_t n, const uint32_t* points,
                          const uint8_t* pinf, const uint32_t* scalars, uint8_t* status,
                          uint8_t* sig_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_sign_finalize, GRID(n, 64), 0, s, jobs, n, points, pinf, scalars, status,
                     sig_out);
}

// ------------------------------------------------------------------ challenge hashing
// c_arr = I2OSP(R, 8) || (I2OSP(i, 8) || m_i)* || Abar || Bbar || D || T1 || T2 ||
//         domain || I2OSP(len(ph), 8) || ph   (proof.rs:2231-2297)
DEVN void challenge_scalar(Fr& out, uint32_t R, const uint32_t* disc_index,
                           const uint32_t* disc_scalar, const uint32_t* scalars,
                           const uint8_t* const pts[5], const Fr& dom, const uint8_t* ph,
                           uint32_t ph_len) {
  Sha256 s;
  sha_init_after_zpad(s);
  sha_u64_be(s, R);
  for (uint32_t r = 0; r < R; r++) {
    sha_u64_be(s, disc_index[r]);
    Fr m;
    load_fr(m, scalars + (size_t)disc_scalar[r] * FR_WORDS);
    sha_scalar(s, m);
  }
  for (int k = 0; k < 5; k++) sha_bytes48(s, pts[k]);
  sha_scalar(s, dom);
  sha_u64_be(s, ph_len);
  sha_bytes_global(s, ph, ph_len);
  xmd_b0_suffix(s, 48, DST_H2S, DST_H2S_LEN);
  uint32_t okm[12];
  xmd48_from_b0(okm, s.h, DST_H2S, DST_H2S_LEN);
  scalar_from_okm48(out, okm);
}

DEVN void compress_point(uint8_t* out, const uint32_t* points, const uint8_t* pinf, uint32_t idx) {
  if (pinf[idx]) {
    out[0] = 0xc0;
    for (int k = 1; k < 48; k++) out[k] = 0;
    return;
  }
  G1A a;
  load_g1a(a, points + (size_t)idx * G1A_WORDS);
  g1_compress_affine(out, a);
}

// ProofVerify: gate[item] = (challenge recomputed from T1, T2 == proof's c)
__global__ __launch_bounds__(64) void k_pv_challenge(const PvChallengeJob* __restrict__ jobs,
                                                    uint32_t n, const uint8_t* __restrict__ blob,
                                                    const uint32_t* __restrict__ disc_index,
                                                    const uint32_t* __restrict__ disc_scalar,
                                                    const uint32_t* __restrict__ points,
                                                    const uint8_t* __restrict__ pinf,
                                                    const uint32_t* __restrict__ scalars,
                                                    uint8_t* __restrict__ gate) {
  wave_priority<PRIO_MID>();
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PvChallengeJob j = jobs[i];
  uint8_t c[5][48];
  compress_point(c[0], points, pinf, j.abar);
  compress_point(c[1], points, pinf, j.bbar);
  compress_point(c[2], points, pinf, j.d);
  compress_point(c[3], points, pinf, j.t1);
  compress_point(c[4], points, pinf, j.t2);
  const uint8_t* pts[5] = {c[0], c[1], c[2], c[3], c[4]};
  Fr dom, chal, want;
  load_fr(dom, scalars + (size_t)j.dom_index * FR_WORDS);
  challenge_scalar(chal, j.R, disc_index + j.disc_begin, disc_scalar + j.disc_begin, scalars, pts,
                   dom, blob + j.ph_off, j.ph_len);
  load_fr(want, scalars + (size_t)j.c_index * FR_WORDS);
  gate[j.item] = limbs_eq<8>(chal.v, want.v) ? 1 : 0;
}
void launch_pv_challenge(hipStream_t s, const PvChallengeJob* jobs, uint32_t n,
                         const uint8_t* blob, const uint32_t* disc_index,
                         const uint32_t* disc_scalar, const uint32_t* points, const uint8_t* pinf,
                         const uint32_t* scalars, uint8_t* gate) {
  if (!n) return;
  hipLaunchKernelGGL(k_pv_challenge, GRID(n, 64), 0, s, jobs, n, blob, disc_index, disc_scalar,
                     points, pinf, scalars, gate);
}

// ------------------------------------------------------------------ ProofGen
DEV uint32_t rotl32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
#define CHACHA_QR(a, b, c, d)                     \
  a += b; d ^= a; d = rotl32(d, 16);              \
  c += d; b ^= c; b = rotl32(b, 12);              \
  a += b; d ^= a; d = rotl32(d, 8);               \
  c += d; b ^= c; b = rotl32(b, 7);

// RFC 8439 block with a zero nonce; out[16] little-endian keystream words.
DEVN void chacha20_block_words(uint32_t out[16], const uint32_t key[8], uint32_t counter) {
  uint32_t st[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u,
                     key[0], key[1], key[2], key[3], key[4], key[5], key[6], key[7],
                     counter, 0u, 0u, 0u};
  uint32_t x[16];
#pragma unroll
  for (int k = 0; k < 16; k++) x[k] = st[k];
#pragma unroll 1
  for (int r = 0; r < 10; r++) {
    CHACHA_QR(x[0], x[4], x[8], x[12]);
    CHACHA_QR(x[1], x[5], x[9], x[13]);
    CHACHA_QR(x[2], x[6], x[10], x[14]);
    CHACHA_QR(x[3], x[7], x[11], x[15]);
    CHACHA_QR(x[0], x[5], x[10], x[15]);
    CHACHA_QR(x[1], x[6], x[11], x[12]);
    CHACHA_QR(x[2], x[7], x[8], x[13]);
    CHACHA_QR(x[3], x[4], x[9], x[14]);
  }
#pragma unroll
  for (int k = 0; k < 16; k++) out[k] = x[k] + st[k];
}

// Blinding scalars from the item's 32-byte seed: scalar k = LE(block k) mod r.
__global__ __launch_bounds__(64) void k_pg_chacha(const PgJob* __restrict__ jobs, uint32_t n,
                                                 const uint8_t* __restrict__ seeds,
                                                 uint32_t* __restrict__ scalars) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PgJob j = jobs[i];
  uint32_t key[8];
  const uint8_t* sd = seeds + (size_t)32 * i;
  for (int k = 0; k < 8; k++)
    key[k] = (uint32_t)sd[4 * k] | ((uint32_t)sd[4 * k + 1] << 8) | ((uint32_t)sd[4 * k + 2] << 16) |
             ((uint32_t)sd[4 * k + 3] << 24);
  for (uint32_t t = 0; t < 5 + j.U; t++) {
    uint32_t w[16];
    chacha20_block_words(w, key, t);
    Fr r;
    fr_reduce_wide(r, w);
    store_fr(scalars + (size_t)(j.rs + t) * FR_WORDS, r);
  }
}
void launch_pg_chacha(hipStream_t s, const PgJob* jobs, uint32_t n, const uint8_t* seeds,
                      uint32_t* scalars) {
  if (!n) return;
  hipLaunchKernelGGL(k_pg_chacha, GRID(n, 64), 0, s, jobs, n, seeds, scalars);
}

// The seven multipliers of A and B:
//   k0 = r2 (D), k1 = r1 r2 (Abar), k2 = r1 r2, k3 = -r1 r2 e (Bbar),
//   k4 = r1 r2 e~, k5 = r2 r1~ (T1), k6 = r2 r3~ (T2 part)
__global__ __launch_bounds__(64) void k_pg_scalars(const PgJob* __restrict__ jobs, uint32_t n,
                                                  uint32_t* __restrict__ scalars,
                                                  uint8_t* __restrict__ status) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PgJob j = jobs[i];
  Fr r1, r2, et, r1t, r3t, e;
  load_fr(r1, scalars + (size_t)(j.rs + 0) * FR_WORDS);
  load_fr(r2, scalars + (size_t)(j.rs + 1) * FR_WORDS);
  load_fr(et, scalars + (size_t)(j.rs + 2) * FR_WORDS);
  load_fr(r1t, scalars + (size_t)(j.rs + 3) * FR_WORDS);
  load_fr(r3t, scalars + (size_t)(j.rs + 4) * FR_WORDS);
  load_fr(e, scalars + (size_t)j.e_index * FR_WORDS);
  Fr r2m, r1r2, r1r2m, k3, k4, k5, k6, t;
  fr_to_mont(r2m, r2);
  fr_mul(r1r2, r1, r2m);        // canonical r1 r2
  fr_to_mont(r1r2m, r1r2);
  fr_mul(t, e, r1r2m);          // r1 r2 e
  fr_neg(k3, t);
  fr_mul(k4, et, r1r2m);
  fr_mul(k5, r1t, r2m);
  fr_mul(k6, r3t, r2m);
  uint32_t* o = scalars + (size_t)j.k_index * FR_WORDS;
  store_fr(o + 0 * FR_WORDS, r2);
  store_fr(o + 1 * FR_WORDS, r1r2);
  store_fr(o + 2 * FR_WORDS, r1r2);
  store_fr(o + 3 * FR_WORDS, k3);
  store_fr(o + 4 * FR_WORDS, k4);
  store_fr(o + 5 * FR_WORDS, k5);
  store_fr(o + 6 * FR_WORDS, k6);
  if (fr_is_zero(r2)) status[j.item] = 1;  // r2 must be invertible (proof.rs:2336-2337)
}
void launch_pg_scalars(hipStream_t s, const PgJob* jobs, uint32_t n, uint32_t* scalars,
                       uint8_t* status) {
  if (!n) return;
  hipLaunchKernelGGL(k_pg_scalars, GRID(n, 64), 0, s, jobs, n, scalars, status);
}

// Challenge and proof_finalize_parts (proof.rs:2317-2377); writes
// Abar || Bbar || D || e^ || r1^ || r3^ || m^_1..m^_U || c.
__global__ __launch_bounds__(64) void k_pg_finalize(const PgJob* __restrict__ jobs, uint32_t n,
                                                   const uint8_t* __restrict__ blob,
                                                   const uint32_t* __restrict__ disc_index,
                                                   const uint32_t* __restrict__ disc_scalar,
                                                   const uint32_t* __restrict__ und_scalar,
                                                   const uint32_t* __restrict__ points,
                                                   const uint8_t* __restrict__ pinf,
                                                   uint32_t* __restrict__ scalars,
                                                   uint8_t* __restrict__ status,
                                                   uint8_t* __restrict__ proof_out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  PgJob j = jobs[i];
  if (status[j.item]) return;
  uint8_t c[5][48];
  compress_point(c[0], points, pinf, j.abar);
  compress_point(c[1], points, pinf, j.bbar);
  compress_point(c[2], points, pinf, j.d);
  compress_point(c[3], points, pinf, j.t1);
  compress_point(c[4], points, pinf, j.t2);
  const uint8_t* pts[5] = {c[0], c[1], c[2], c[3], c[4]};
  Fr dom, ch;
  load_fr(dom, scalars + (size_t)j.dom_index * FR_WORDS);
  challenge_scalar(ch, j.R, disc_index + j.disc_begin, disc_scalar + j.disc_begin, scalars, pts,
                   dom, blob + j.ph_off, j.ph_len);
  Fr chm, r1, r2, et, r1t, r3t, e, t, x;
  fr_to_mont(chm, ch);
  load_fr(r1, scalars + (size_t)(j.rs + 0) * FR_WORDS);
  load_fr(r2, scalars + (size_t)(j.rs + 1) * FR_WORDS);
  load_fr(et, scalars + (size_t)(j.rs + 2) * FR_WORDS);
  load_fr(r1t, scalars + (size_t)(j.rs + 3) * FR_WORDS);
  load_fr(r3t, scalars + (size_t)(j.rs + 4) * FR_WORDS);
  load_fr(e, scalars + (size_t)j.e_index * FR_WORDS);
  uint8_t* o = proof_out + j.proof_off;
  for (int k = 0; k < 48; k++) {
    o[k] = c[0][k];
    o[48 + k] = c[1][k];
    o[96 + k] = c[2][k];
  }
  // e^ = e~ + e c
  fr_mul(t, e, chm);
  fr_add(x, et, t);
  fr_to_be_bytes(o + 144, x);
  // r1^ = r1~ - r1 c
  fr_mul(t, r1, chm);
  fr_sub(x, r1t, t);
  fr_to_be_bytes(o + 176, x);
  // r3^ = r3~ - r2^-1 c
  Fr r2m, r3;
  fr_to_mont(r2m, r2);
  fr_inv(r3, r2m);         // Montgomery r2^-1
  fr_mul(t, chm, r3);      // Montgomery (c r2^-1) -> from_mont
  fr_from_mont(t, t);
  fr_sub(x, r3t, t);
  fr_to_be_bytes(o + 208, x);
  // m^_j = m~_j + m_j c
  for (uint32_t u = 0; u < j.U; u++) {
    Fr mt, m;
    load_fr(mt, scalars + (size_t)(j.rs + 5 + u) * FR_WORDS);
    load_fr(m, scalars + (size_t)und_scalar[j.und_begin + u] * FR_WORDS);
    fr_mul(t, m, chm);
    fr_add(x, mt, t);
    fr_to_be_bytes(o + 240 + 32 * u, x);
  }
  fr_to_be_bytes(o + 240 + 32 * j.U, ch);
}
void launch_pg_finalize(hipStream_t s, const PgJob* jobs, uint32_t n, const uint8_t* blob,
                        const uint32_t* disc_index, const uint32_t* disc_scalar,
                        const uint32_t* und_scalar, const uint32_t* points, const uint8_t* pinf,
                        uint32_t* scalars, uint8_t* status, uint8_t* proof_out) {
  if (!n) return;
  hipLaunchKernelGGL(k_pg_finalize, GRID(n, 64), 0, s, jobs, n, blob, disc_index, disc_scalar,
                     und_scalar, points, pinf, scalars, status, proof_out);
}

// ------------------------------------------------------------------ keys
// PK = sk * G2 (crates/cubbs-bbs-protocol/src/bbsplus/keys.rs:256-260)
__global__ __launch_bounds__(64) void k_g2_mul_gen(const uint8_t* sk, uint32_t n, uint8_t* pk_out,
                                                  uint8_t* ok) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Fr k;
  fr_from_be_bytes_raw(k, sk + (size_t)32 * i);
  if (!fr_canonical(k)) {
    ok[i] = 0;
    return;
  }
  G2A g;
  g2_load_affine_const(g, bls::G2_GEN);
  G2J r;
  pt_mul_limbs_affine<Fp2, 8>(r, g, k.v);
  uint8_t c[96];
  g2_compress(c, r);
  for (int q = 0; q < 96; q++) pk_out[(size_t)96 * i + q] = c[q];
  ok[i] = 1;
}
void launch_g2_mul_gen(hipStream_t s, const uint8_t* sk, uint32_t n, uint8_t* pk_out, uint8_t* ok) {
  if (!n) return;
  hipLaunchKernelGGL(k_g2_mul_gen, GRID(n, 64), 0, s, sk, n, pk_out, ok);
}

}  // namespace cubbs
