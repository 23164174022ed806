// Load/store helpers between flat device arrays (32-bit words) and the
// field / point types; the issue priority of pipeline stages.
#pragma once
#include "pairing.hpp"
#include "sha.hpp"

// Issue priority of a kernel's waves (s_setprio, within the SIMD arbiter;
// streams keep the default priority).  With many batches in flight, the
// stages at the start of a batch's chain go first: decode, hashing, term
// fills and the variable-base products at PRIO_EARLY, the comb, the final
// sums and the challenge at PRIO_MID, the pairing stages at the default 0.
// The batches then move through their chains together, and their last
// stages overlap instead of draining one batch after another (twenty
// Verify+ProofVerify sets: 2.55-2.65 ms per step against 2.62-2.68).  The
// Pippenger stages use PRIO_PIP, above the subgroup checks that run beside
// them in the product kernel.
constexpr int PRIO_MID = 1, PRIO_EARLY = 2, PRIO_PIP = 3;
template <int P>
DEV void wave_priority() {
  __builtin_amdgcn_s_setprio(P);
}

DEV void load_fr(Fr& r, const uint32_t* p) {
#pragma unroll
  for (int k = 0; k < 8; k++) r.v[k] = p[k];
}
DEV void store_fr(uint32_t* p, const Fr& r) {
#pragma unroll
  for (int k = 0; k < 8; k++) p[k] = r.v[k];
}
DEV void ld_fp(Fp& r, const uint32_t* p) {
#pragma unroll
  for (int k = 0; k < FPW; k++) r.v[k] = p[k];
}
DEV void st_fp(uint32_t* p, const Fp& v) {
#pragma unroll
  for (int k = 0; k < FPW; k++) p[k] = v.v[k];
}
DEV void load_g1a(G1A& r, const uint32_t* p) {
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    r.x.v[k] = p[k];
    r.y.v[k] = p[FPW + k];
  }
}
DEV void store_g1a(uint32_t* p, const G1A& r) {
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    p[k] = r.x.v[k];
    p[FPW + k] = r.y.v[k];
  }
}
DEV void load_g1j(G1J& r, const uint32_t* p) {
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    r.x.v[k] = p[k];
    r.y.v[k] = p[FPW + k];
    r.z.v[k] = p[2 * FPW + k];
  }
}
DEV void store_g1j(uint32_t* p, const G1J& r) {
#pragma unroll
  for (int k = 0; k < FPW; k++) {
    p[k] = r.x.v[k];
    p[FPW + k] = r.y.v[k];
    p[2 * FPW + k] = r.z.v[k];
  }
}
// 32-byte big-endian scalar as 8 SHA words (most significant first)
DEV void sha_scalar(Sha256& s, const Fr& canonical) {
#pragma unroll
  for (int k = 7; k >= 0; k--) sha_word(s, canonical.v[k]);
}
DEV void sha_bytes48(Sha256& s, const uint8_t* c) {
#pragma unroll
  for (int k = 0; k < 12; k++)
    sha_word(s, ((uint32_t)c[4 * k] << 24) | ((uint32_t)c[4 * k + 1] << 16) |
                    ((uint32_t)c[4 * k + 2] << 8) | c[4 * k + 3]);
}
