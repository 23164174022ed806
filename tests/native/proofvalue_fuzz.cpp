// Seeded mutation loop over the proofValue codec's C entries, built with
// g++ -fsanitize=address,undefined by tests/test_proofvalue.py (host code
// only).  Valid base and derived proofValues are mutated (byte overwrites,
// bit flips, truncations, insertions, stray characters, short arenas and
// misaligned arena starts) and parsed; a sanitizer report or a wrong result
// on the unmutated inputs fails the run.  Prints the count of each result
// code.
#include "cubbs_proofvalue.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

static std::vector<uint8_t> unb64(const char* s, size_t n) {
  static const char A[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789-_";
  std::vector<uint8_t> o;
  uint32_t acc = 0;
  int bits = 0;
  for (size_t i = 1; i < n; i++) {  // skips the multibase 'u'
    acc = acc << 6 | (uint32_t)(strchr(A, s[i]) - A);
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      o.push_back((uint8_t)(acc >> bits));
      acc &= (1u << bits) - 1;
    }
  }
  return o;
}

static std::string b64(const std::vector<uint8_t>& b) {
  static const char A[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789-_";
  std::string o = "u";
  size_t i = 0;
  for (; i + 3 <= b.size(); i += 3) {
    const uint32_t v = b[i] << 16 | b[i + 1] << 8 | b[i + 2];
    o += A[v >> 18];
    o += A[(v >> 12) & 63];
    o += A[(v >> 6) & 63];
    o += A[v & 63];
  }
  if (b.size() - i == 1) {
    const uint32_t v = b[i] << 16;
    o += A[v >> 18];
    o += A[(v >> 12) & 63];
  } else if (b.size() - i == 2) {
    const uint32_t v = b[i] << 16 | b[i + 1] << 8;
    o += A[v >> 18];
    o += A[(v >> 12) & 63];
    o += A[(v >> 6) & 63];
  }
  return o;
}

int main(int argc, char** argv) {
  const long iters = argc > 1 ? atol(argv[1]) : 100000;
  std::mt19937_64 rng(3);
  std::vector<uint8_t> proof(300, 7), ph(10, 3), sig(80, 1), hd(64, 2), pk(96, 3), hm(32, 4);
  const char* ks[5] = {"c14n0", "c14n1", "c14n22", "c14n3", "c14n+4"};
  const char* vs[5] = {"b3", "b2", "b1", "b0", "b9"};
  cubbs_pv_label lab[5];
  for (int i = 0; i < 5; i++) {
    lab[i].key = {(const uint8_t*)ks[i], strlen(ks[i])};
    lab[i].value = {(const uint8_t*)vs[i], strlen(vs[i])};
  }
  uint64_t man[4] = {0, 1, 2, 1ull << 40}, sel[3] = {9, 14, 300};
  cubbs_pv_derived d{{proof.data(), proof.size()}, lab, 5, man, 4, sel, 3, {ph.data(), ph.size()}};
  const char* p1 = "/issuer";
  const char* p2 = "/credentialSubject/x";
  cubbs_pv_slice ptrs[2] = {{(const uint8_t*)p1, strlen(p1)}, {(const uint8_t*)p2, strlen(p2)}};
  cubbs_pv_base b{{sig.data(), 80}, {hd.data(), 64}, {pk.data(), 96}, {hm.data(), 32}, ptrs, 2};
  std::vector<char> out(8192);
  size_t n1 = 0, n2 = 0;
  if (cubbs_proofvalue_serialize_derived(&d, out.data(), out.size(), &n1) != 0) return 2;
  const std::vector<uint8_t> rd = unb64(out.data(), n1);
  if (cubbs_proofvalue_serialize_base(&b, out.data(), out.size(), &n2) != 0) return 2;
  const std::vector<uint8_t> rb = unb64(out.data(), n2);
  std::vector<char> arena(1 << 16);
  size_t need;
  cubbs_pv_derived od;
  cubbs_pv_base ob;
  // the unmutated values parse back
  std::string s = b64(rd);
  if (cubbs_proofvalue_parse_derived(s.data(), s.size(), arena.data(), arena.size(), &need, &od) != 0 ||
      od.bbs_proof.len != 300 || od.n_labels != 5 || od.n_selective_indexes != 3)
    return 3;
  s = b64(rb);
  if (cubbs_proofvalue_parse_base(s.data(), s.size(), arena.data(), arena.size(), &need, &ob) != 0 ||
      ob.n_mandatory_pointers != 2 || ob.public_key.len != 96)
    return 3;
  long counts[10] = {0};
  for (long it = 0; it < iters; it++) {
    std::vector<uint8_t> m = (it & 1) ? rd : rb;
    const int k = 1 + (int)(rng() % 4);
    for (int j = 0; j < k && m.size() > 4; j++) {
      const size_t pos = 3 + rng() % (m.size() - 3);
      switch (rng() % 4) {
        case 0: m[pos] = (uint8_t)rng(); break;
        case 1: m[pos] ^= (uint8_t)(1u << (rng() % 8)); break;
        case 2: m.resize(pos); break;
        default: m.insert(m.begin() + pos, (uint8_t)rng()); break;
      }
    }
    s = b64(m);
    if (rng() % 50 == 0 && s.size() > 2) s[1 + rng() % (s.size() - 1)] = (char)rng();
    if (rng() % 50 == 0) s.resize(rng() % s.size());
    const size_t cap = (rng() % 8 == 0) ? rng() % 2000 : arena.size() - 8;
    char* a = arena.data() + rng() % 8;
    const int rc = (it & 1) ? cubbs_proofvalue_parse_derived(s.data(), s.size(), a, cap, &need, &od)
                            : cubbs_proofvalue_parse_base(s.data(), s.size(), a, cap, &need, &ob);
    if (rc < 0 || rc > 9) return 4;
    counts[rc]++;
    if (rc == 0 && (it & 1)) {  // touch every output byte
      volatile uint8_t x = 0;
      for (size_t q = 0; q < od.bbs_proof.len; q++) x ^= od.bbs_proof.ptr[q];
      for (size_t q = 0; q < od.n_labels; q++) x ^= od.label_map[q].key.ptr[od.label_map[q].key.len - 1];
      for (size_t q = 0; q < od.n_selective_indexes; q++) x ^= (uint8_t)od.selective_indexes[q];
    }
  }
  for (int i = 0; i < 10; i++) printf("%s%ld", i ? " " : "", counts[i]);
  printf("\n");
  return 0;
}
