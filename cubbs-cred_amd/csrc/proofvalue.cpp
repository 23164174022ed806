// W3C `bbs-2023` proofValue codec (host code; no GPU work).
//
// Follows crates/cubbs-bbs-proofvalue/src/lib.rs, whose encoding is
// serde_cbor 0.11.2 over serde_bytes 0.11.19 payload tuples, wrapped by
// base64 0.22.1 URL_SAFE_NO_PAD (versions pinned in the reference's
// Cargo.lock).  The decoder here accepts exactly the CBOR those crates accept
// for the two payload types:
//   * definite and indefinite lengths, and integers in any width (non-minimal
//     forms too); semantic tags in front of any value are skipped;
//   * a bytes field also takes a text string (its UTF-8 bytes) or an array
//     of integers 0..255; a text field also takes a byte string that is
//     valid UTF-8; an integer field takes major type 0 only;
//   * arrays, maps and tags nest at most 127 deep (serde_cbor's recursion
//     limit of 128);
//   * the payload array holds exactly five items, later duplicate map keys
//     replace earlier ones, and nothing may follow the payload.
// The encoder writes serde_cbor's form: definite lengths, minimal integer
// heads, map keys ascending.
#include "../../include/cubbs_proofvalue.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr uint8_t PREFIX_BASE[3] = {0xd9, 0x5d, 0x02};     // lib.rs:11
constexpr uint8_t PREFIX_DERIVED[3] = {0xd9, 0x5d, 0x03};  // lib.rs:12

// ------------------------------------------------------------ base64url, no padding
constexpr char B64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789-_";

// symbol -> 6-bit value, -1 outside the url-safe alphabet
struct B64Table {
  int8_t v[256];
  constexpr B64Table() : v() {
    for (int i = 0; i < 256; i++) v[i] = -1;
    for (int i = 0; i < 64; i++) v[(uint8_t)B64[i]] = (int8_t)i;
  }
};
constexpr B64Table B64T{};

size_t b64_len(size_t n) { return n / 3 * 4 + (n % 3 ? n % 3 + 1 : 0); }

void b64_encode(const uint8_t* p, size_t n, char* o) {
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8 | p[i + 2];
    *o++ = B64[v >> 18];
    *o++ = B64[(v >> 12) & 63];
    *o++ = B64[(v >> 6) & 63];
    *o++ = B64[v & 63];
  }
  if (n - i == 1) {
    const uint32_t v = (uint32_t)p[i] << 16;
    *o++ = B64[v >> 18];
    *o++ = B64[(v >> 12) & 63];
  } else if (n - i == 2) {
    const uint32_t v = (uint32_t)p[i] << 16 | (uint32_t)p[i + 1] << 8;
    *o++ = B64[v >> 18];
    *o++ = B64[(v >> 12) & 63];
    *o++ = B64[(v >> 6) & 63];
  }
}

// base64 0.22 URL_SAFE_NO_PAD decoding: alphabet symbols only (no '='), no
// lone final symbol, and the unused low bits of the last symbol zero.
bool b64_decode(const uint8_t* s, size_t n, std::vector<uint8_t>& out) {
  if (n % 4 == 1) return false;
  out.resize(n / 4 * 3 + (n % 4 ? n % 4 - 1 : 0));
  uint8_t* o = out.data();
  const int8_t* T = B64T.v;
  size_t i = 0;
  for (; i + 4 <= n; i += 4, o += 3) {
    const int a = T[s[i]], b = T[s[i + 1]], c = T[s[i + 2]], d = T[s[i + 3]];
    if ((a | b | c | d) < 0) return false;
    const uint32_t v = (uint32_t)a << 18 | (uint32_t)b << 12 | (uint32_t)c << 6 | (uint32_t)d;
    o[0] = (uint8_t)(v >> 16);
    o[1] = (uint8_t)(v >> 8);
    o[2] = (uint8_t)v;
  }
  if (n - i == 2) {  // 12 bits: one byte and 4 spare bits
    const int a = T[s[i]], b = T[s[i + 1]];
    if ((a | b) < 0 || (b & 15)) return false;
    o[0] = (uint8_t)(a << 2 | b >> 4);
  } else if (n - i == 3) {  // 18 bits: two bytes and 2 spare bits
    const int a = T[s[i]], b = T[s[i + 1]], c = T[s[i + 2]];
    if ((a | b | c) < 0 || (c & 3)) return false;
    const uint32_t v = (uint32_t)a << 12 | (uint32_t)b << 6 | (uint32_t)c;
    o[0] = (uint8_t)(v >> 10);
    o[1] = (uint8_t)(v >> 2);
  }
  return true;
}

// ------------------------------------------------------------ UTF-8 (Rust's str::from_utf8)
bool utf8_ok(const uint8_t* p, size_t n) {
  size_t i = 0;
  while (i < n) {
    const uint8_t c = p[i];
    if (c < 0x80) {
      i++;
      continue;
    }
    int need;
    uint8_t lo = 0x80, hi = 0xbf;
    if (c >= 0xc2 && c <= 0xdf) {
      need = 1;
    } else if (c >= 0xe0 && c <= 0xef) {
      need = 2;
      if (c == 0xe0) lo = 0xa0;        // no overlong forms
      else if (c == 0xed) hi = 0x9f;   // no surrogates
    } else if (c >= 0xf0 && c <= 0xf4) {
      need = 3;
      if (c == 0xf0) lo = 0x90;        // no overlong forms
      else if (c == 0xf4) hi = 0x8f;   // at most U+10FFFF
    } else {
      return false;
    }
    if (n - i - 1 < (size_t)need) return false;
    if (p[i + 1] < lo || p[i + 1] > hi) return false;
    for (int k = 2; k <= need; k++)
      if (p[i + k] < 0x80 || p[i + k] > 0xbf) return false;
    i += need + 1;
  }
  return true;
}

// ------------------------------------------------------------ CBOR writer (serde_cbor's form)
struct Writer {
  std::vector<uint8_t> b;
  void head(uint8_t major, uint64_t v) {
    const uint8_t m = (uint8_t)(major << 5);
    if (v < 24) {
      b.push_back(m | (uint8_t)v);
    } else if (v <= 0xff) {
      b.push_back(m | 24);
      b.push_back((uint8_t)v);
    } else if (v <= 0xffff) {
      b.push_back(m | 25);
      for (int s = 8; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    } else if (v <= 0xffffffffull) {
      b.push_back(m | 26);
      for (int s = 24; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    } else {
      b.push_back(m | 27);
      for (int s = 56; s >= 0; s -= 8) b.push_back((uint8_t)(v >> s));
    }
  }
  void run(uint8_t major, const uint8_t* p, size_t n) {
    head(major, n);
    if (n) b.insert(b.end(), p, p + n);
  }
};

// ------------------------------------------------------------ CBOR reader (serde_cbor + serde visitors)
// A decoded byte/text value: a view into the input, or assembled bytes.
struct Run {
  const uint8_t* p = nullptr;
  size_t n = 0;
  bool owned = false;
  std::string own;
  void view(const uint8_t* q, size_t len) {
    p = q;
    n = len;
    owned = false;
  }
  const uint8_t* data() const { return owned ? (const uint8_t*)own.data() : p; }
  size_t size() const { return owned ? own.size() : n; }
};

struct Reader {
  const uint8_t* p;
  size_t n;
  size_t i = 0;
  int depth = 128;  // serde_cbor remaining_depth

  bool byte(uint8_t& c) {
    if (i >= n) return false;
    c = p[i++];
    return true;
  }
  bool peek(uint8_t& c) const {
    if (i >= n) return false;
    c = p[i];
    return true;
  }
  // Argument of an initial byte with additional info ai < 28.
  bool arg(uint8_t ai, uint64_t& v) {
    if (ai < 24) {
      v = ai;
      return true;
    }
    if (ai > 27) return false;
    const int len = 1 << (ai - 24);
    if (n - i < (size_t)len) return false;
    v = 0;
    for (int k = 0; k < len; k++) v = v << 8 | p[i++];
    return true;
  }
  bool enter() { return --depth > 0; }
  void leave() { depth++; }
  // Reads the head of the next value, skipping semantic tags (each tag is
  // one nesting level until the tagged value ends, as in serde_cbor).
  // Returns the number of levels entered in *tags.
  bool head(uint8_t& major, uint8_t& ai, int& tags) {
    tags = 0;
    for (;;) {
      uint8_t c;
      if (!byte(c)) return false;
      major = c >> 5;
      ai = c & 31;
      if (major != 6) return true;
      uint64_t tag;
      if (!arg(ai, tag)) return false;  // 0xdc..0xdf unassigned
      if (!enter()) return false;
      tags++;
    }
  }
  void untag(int tags) { depth += tags; }
  // A definite run of len bytes.
  bool take(uint64_t len, const uint8_t*& q) {
    if (len > n - i) return false;
    q = p + i;
    i += (size_t)len;
    return true;
  }
  // Body of a byte (major 2) or text (major 3) string after its head;
  // indefinite strings are chunks of the same major type up to a break.
  // A definite string stays a view into the input.
  bool string_body(uint8_t major, uint8_t ai, Run& out) {
    if (ai == 31) {
      out.own.clear();
      out.owned = true;
      for (;;) {
        uint8_t c;
        if (!byte(c)) return false;
        if (c == 0xff) break;
        if ((c >> 5) != major || (c & 31) == 31) return false;
        uint64_t len;
        const uint8_t* q;
        if (!arg(c & 31, len) || !take(len, q)) return false;
        out.own.append((const char*)q, (size_t)len);
      }
    } else {
      uint64_t len;
      const uint8_t* q;
      if (!arg(ai, len) || !take(len, q)) return false;
      out.view(q, (size_t)len);
    }
    return major == 2 || utf8_ok(out.data(), out.size());
  }
  // Visits a sequence (major 4) after its head: elem() is called for each
  // item; a definite sequence must be consumed whole.
  template <class F>
  bool seq_body(uint8_t ai, F&& elem) {
    if (!enter()) return false;
    if (ai == 31) {
      for (;;) {
        uint8_t c;
        if (!peek(c)) return false;
        if (c == 0xff) {
          i++;
          break;
        }
        if (!elem()) return false;
      }
    } else {
      uint64_t len;
      if (!arg(ai, len)) return false;
      for (uint64_t k = 0; k < len; k++)
        if (!elem()) return false;
    }
    leave();
    return true;
  }

  bool u64v(uint64_t& v) {
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags) || major != 0 || !arg(ai, v)) return false;
    untag(tags);
    return true;
  }
  // serde_bytes ByteBuf: bytes, text, or an array of u8
  bool bytesv(Run& out) {
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags)) return false;
    bool ok;
    if (major == 2 || major == 3) {
      ok = string_body(major, ai, out);
    } else if (major == 4) {
      out.own.clear();
      out.owned = true;
      ok = seq_body(ai, [&] {
        uint64_t v;
        if (!u64v(v) || v > 255) return false;
        out.own.push_back((char)v);
        return true;
      });
    } else {
      ok = false;
    }
    if (ok) untag(tags);
    return ok;
  }
  // serde String: text, or bytes that are UTF-8
  bool textv(Run& out) {
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags) || (major != 2 && major != 3)) return false;
    if (!string_body(major, ai, out) || !utf8_ok(out.data(), out.size())) return false;
    untag(tags);
    return true;
  }
  template <class F>
  bool seqv(F&& elem) {
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags) || major != 4 || !seq_body(ai, elem)) return false;
    untag(tags);
    return true;
  }
  bool u64_listv(std::vector<uint64_t>& out) {
    out.clear();
    return seqv([&] {
      uint64_t v;
      if (!u64v(v)) return false;
      out.push_back(v);
      return true;
    });
  }
  // entries in stream order; the caller keeps the last of equal keys
  bool u64_mapv(std::vector<std::pair<uint64_t, uint64_t>>& out) {
    out.clear();
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags) || major != 5 || !enter()) return false;
    auto entry = [&] {
      uint64_t k, v;
      if (!u64v(k) || !u64v(v)) return false;
      out.emplace_back(k, v);
      return true;
    };
    if (ai == 31) {
      for (;;) {
        uint8_t c;
        if (!peek(c)) return false;
        if (c == 0xff) {
          i++;
          break;
        }
        if (!entry()) return false;
      }
    } else {
      uint64_t len;
      if (!arg(ai, len)) return false;
      for (uint64_t k = 0; k < len; k++)
        if (!entry()) return false;
    }
    leave();
    untag(tags);
    return true;
  }
  // The payload tuple struct: an array of exactly `count` items.
  template <class F>
  bool tuplev(int count, F&& field) {
    uint8_t major, ai;
    int tags;
    if (!head(major, ai, tags) || major != 4 || !enter()) return false;
    if (ai == 31) {
      for (int k = 0; k < count; k++) {
        uint8_t c;
        if (!peek(c) || c == 0xff || !field(k)) return false;
      }
      uint8_t c;
      if (!byte(c) || c != 0xff) return false;
    } else {
      uint64_t len;
      if (!arg(ai, len) || len != (uint64_t)count) return false;
      for (int k = 0; k < count; k++)
        if (!field(k)) return false;
    }
    leave();
    untag(tags);
    return i == n;  // from_slice: nothing may follow
  }
};

// ------------------------------------------------------------ encoding
// lib.rs:145-153
bool valid_base(size_t sig, size_t header, size_t pk, size_t hmac) {
  return sig == 80 && header == 64 && pk == 96 && hmac == 32;
}

// Rust's u64::from_str: an optional '+', then one or more ASCII digits, no
// overflow.
bool parse_u64(const uint8_t* s, size_t n, uint64_t& v) {
  if (n && s[0] == '+') {
    s++;
    n--;
  }
  if (!n) return false;
  v = 0;
  for (size_t k = 0; k < n; k++) {
    if (s[k] < '0' || s[k] > '9') return false;
    const uint64_t d = s[k] - '0';
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  return true;
}

// std::string (and Rust String) order on byte runs
int run_cmp(const cubbs_pv_slice& a, const cubbs_pv_slice& b) {
  const size_t n = a.len < b.len ? a.len : b.len;
  const int c = n ? memcmp(a.ptr, b.ptr, n) : 0;
  return c ? c : (a.len < b.len ? -1 : a.len > b.len ? 1 : 0);
}

// Sorts (key, value) pairs by key, keeping the last pair given of equal
// keys: the contents of a BTreeMap filled by insert() in the given order.
template <class V, class Less, class Eq>
void btree_insert_order(std::vector<V>& v, Less less, Eq eq) {
  std::stable_sort(v.begin(), v.end(), less);
  size_t n = 0;
  for (size_t k = 0; k < v.size(); k++)
    if (k + 1 == v.size() || !eq(v[k + 1], v[k])) v[n++] = v[k];
  v.resize(n);
}

// Encoder state reused by the calls of one thread.
struct EncodeScratch {
  Writer w;
  std::vector<const cubbs_pv_label*> labels;
  std::vector<std::pair<uint64_t, uint64_t>> comp;
};
thread_local EncodeScratch tes;

int32_t emit(const std::vector<uint8_t>& bytes, char* out, size_t cap, size_t* out_len) {
  const size_t need = 1 + b64_len(bytes.size());
  *out_len = need;
  if (cap < need || !out) return CUBBS_PV_ERR_INSUFFICIENT_OUTPUT;
  out[0] = 'u';
  b64_encode(bytes.data(), bytes.size(), out + 1);
  return CUBBS_PV_OK;
}

// lib.rs:132-143: multibase 'u', base64url, then the 3-byte prefix.
int32_t unwrap(const char* pv, size_t len, const uint8_t* prefix, std::vector<uint8_t>& bytes) {
  if (len == 0 || pv[0] != 'u') return CUBBS_PV_ERR_MULTIBASE_PREFIX;
  if (!b64_decode((const uint8_t*)pv + 1, len - 1, bytes)) return CUBBS_PV_ERR_BASE64;
  if (bytes.size() < 3 || memcmp(bytes.data(), prefix, 3) != 0) return CUBBS_PV_ERR_CBOR_PREFIX;
  return CUBBS_PV_OK;
}

bool slice_ok(const cubbs_pv_slice& s) { return s.ptr || !s.len; }

// Bump allocation inside the caller's arena.
struct Arena {
  uint8_t* base;
  size_t off = 0;
  template <class T>
  T* take(size_t count) {
    off = (off + alignof(T) - 1) / alignof(T) * alignof(T);
    T* r = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return r;
  }
  cubbs_pv_slice copy(const uint8_t* p, size_t n) {
    uint8_t* d = take<uint8_t>(n);
    if (n) memcpy(d, p, n);
    return {d, n};
  }
  cubbs_pv_slice copy(const Run& r) { return copy(r.data(), r.size()); }
};
// Bytes the arena needs for `used` bytes laid out by Arena from an
// arbitrarily aligned start.
size_t arena_size(size_t used) { return used + alignof(uint64_t) - 1; }

uint8_t* aligned(void* arena) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(arena);
  return reinterpret_cast<uint8_t*>((a + alignof(uint64_t) - 1) & ~(uintptr_t)(alignof(uint64_t) - 1));
}

}  // namespace

static int32_t serialize_base_impl(const cubbs_pv_base* value, char* out,
                                                   size_t out_cap, size_t* out_len) {
  if (!value || !out_len) return CUBBS_PV_ERR_ARGUMENT;
  *out_len = 0;
  const cubbs_pv_slice* f[4] = {&value->bbs_signature, &value->bbs_header, &value->public_key,
                                &value->hmac_key};
  for (const cubbs_pv_slice* x : f)
    if (!slice_ok(*x)) return CUBBS_PV_ERR_ARGUMENT;
  if (!value->mandatory_pointers && value->n_mandatory_pointers) return CUBBS_PV_ERR_ARGUMENT;
  for (size_t k = 0; k < value->n_mandatory_pointers; k++) {
    const cubbs_pv_slice& s = value->mandatory_pointers[k];
    if (!slice_ok(s) || !utf8_ok(s.ptr, s.len)) return CUBBS_PV_ERR_ARGUMENT;
  }
  if (!valid_base(f[0]->len, f[1]->len, f[2]->len, f[3]->len)) return CUBBS_PV_ERR_INVALID_BASE;
  // lib.rs:73-83, 123-130
  Writer& w = tes.w;
  w.b.assign(PREFIX_BASE, PREFIX_BASE + 3);
  w.head(4, 5);
  for (const cubbs_pv_slice* x : f) w.run(2, x->ptr, x->len);
  w.head(4, value->n_mandatory_pointers);
  for (size_t k = 0; k < value->n_mandatory_pointers; k++)
    w.run(3, value->mandatory_pointers[k].ptr, value->mandatory_pointers[k].len);
  return emit(w.b, out, out_cap, out_len);
}

static int32_t serialize_derived_impl(const cubbs_pv_derived* value, char* out,
                                                      size_t out_cap, size_t* out_len) {
  if (!value || !out_len) return CUBBS_PV_ERR_ARGUMENT;
  *out_len = 0;
  if (!slice_ok(value->bbs_proof) || !slice_ok(value->presentation_header) ||
      (!value->label_map && value->n_labels) ||
      (!value->mandatory_indexes && value->n_mandatory_indexes) ||
      (!value->selective_indexes && value->n_selective_indexes))
    return CUBBS_PV_ERR_ARGUMENT;
  EncodeScratch& t = tes;
  t.labels.clear();
  for (size_t k = 0; k < value->n_labels; k++) {
    const cubbs_pv_label& l = value->label_map[k];
    if (!slice_ok(l.key) || !slice_ok(l.value) || !utf8_ok(l.key.ptr, l.key.len) ||
        !utf8_ok(l.value.ptr, l.value.len))
      return CUBBS_PV_ERR_ARGUMENT;
    t.labels.push_back(&l);
  }
  if (!value->bbs_proof.len) return CUBBS_PV_ERR_INVALID_DERIVED;  // lib.rs:155-160
  // the input map as a BTreeMap<String, String>
  btree_insert_order(
      t.labels, [](const cubbs_pv_label* a, const cubbs_pv_label* b) { return run_cmp(a->key, b->key) < 0; },
      [](const cubbs_pv_label* a, const cubbs_pv_label* b) { return run_cmp(a->key, b->key) == 0; });
  // lib.rs:165-183: "c14n<u64>" -> "b<u64>", inserted in key order
  t.comp.clear();
  for (const cubbs_pv_label* l : t.labels) {
    uint64_t a, b;
    if (l->key.len < 4 || memcmp(l->key.ptr, "c14n", 4) != 0 || !parse_u64(l->key.ptr + 4, l->key.len - 4, a))
      return CUBBS_PV_ERR_INVALID_DERIVED;
    if (!l->value.len || l->value.ptr[0] != 'b' || !parse_u64(l->value.ptr + 1, l->value.len - 1, b))
      return CUBBS_PV_ERR_INVALID_DERIVED;
    t.comp.emplace_back(a, b);
  }
  btree_insert_order(
      t.comp, [](const auto& a, const auto& b) { return a.first < b.first; },
      [](const auto& a, const auto& b) { return a.first == b.first; });
  // lib.rs:98-108, 123-130
  Writer& w = t.w;
  w.b.assign(PREFIX_DERIVED, PREFIX_DERIVED + 3);
  w.head(4, 5);
  w.run(2, value->bbs_proof.ptr, value->bbs_proof.len);
  w.head(5, t.comp.size());
  for (const auto& kv : t.comp) {
    w.head(0, kv.first);
    w.head(0, kv.second);
  }
  w.head(4, value->n_mandatory_indexes);
  for (size_t k = 0; k < value->n_mandatory_indexes; k++) w.head(0, value->mandatory_indexes[k]);
  w.head(4, value->n_selective_indexes);
  for (size_t k = 0; k < value->n_selective_indexes; k++) w.head(0, value->selective_indexes[k]);
  w.run(2, value->presentation_header.ptr, value->presentation_header.len);
  return emit(w.b, out, out_cap, out_len);
}

namespace {

// Decoded values of one parse call, kept per thread so that repeated calls
// reuse their storage.
struct ParseScratch {
  std::vector<uint8_t> bytes;
  Run f[4];
  std::vector<Run> pointers;
  std::vector<std::pair<uint64_t, uint64_t>> labels;
  std::vector<uint64_t> mandatory, selective;
};
thread_local ParseScratch tls;

// A decompressed label "c14n<k>" -> "b<v>" (lib.rs:185-189).
struct Label {
  char k[24];  // "c14n" + up to 20 digits
  char v[21];  // "b" + up to 20 digits
  uint8_t kl, vl;
};
int decimal(char* o, uint64_t x) {
  char t[20];
  int n = 0;
  do {
    t[n++] = (char)('0' + x % 10);
    x /= 10;
  } while (x);
  for (int k = 0; k < n; k++) o[k] = t[n - 1 - k];
  return n;
}
// std::string order: bytes, then the shorter first
bool label_less(const Label& a, const Label& b) {
  const int c = memcmp(a.k, b.k, a.kl < b.kl ? a.kl : b.kl);
  return c ? c < 0 : a.kl < b.kl;
}

}  // namespace

static int32_t parse_base_impl(const char* proof_value, size_t len, void* arena,
                                               size_t arena_cap, size_t* arena_needed, cubbs_pv_base* out) {
  if (!arena_needed || !out || (!proof_value && len)) return CUBBS_PV_ERR_ARGUMENT;
  *arena_needed = 0;
  ParseScratch& t = tls;
  int32_t rc = unwrap(proof_value, len, PREFIX_BASE, t.bytes);
  if (rc) return rc;
  Reader r{t.bytes.data() + 3, t.bytes.size() - 3};
  size_t np = 0;
  const bool ok = r.tuplev(5, [&](int k) {
    if (k < 4) return r.bytesv(t.f[k]);
    return r.seqv([&] {
      if (np == t.pointers.size()) t.pointers.emplace_back();
      return r.textv(t.pointers[np++]);
    });
  });
  if (!ok) return CUBBS_PV_ERR_CBOR;
  // lib.rs:145-153
  if (t.f[0].size() != 80 || t.f[1].size() != 64 || t.f[2].size() != 96 || t.f[3].size() != 32)
    return CUBBS_PV_ERR_INVALID_BASE;
  size_t used = np * sizeof(cubbs_pv_slice) + 80 + 64 + 96 + 32;
  for (size_t k = 0; k < np; k++) used += t.pointers[k].size();
  *arena_needed = arena_size(used);
  if (!arena || arena_cap < *arena_needed) return CUBBS_PV_ERR_INSUFFICIENT_OUTPUT;
  Arena a{aligned(arena)};
  cubbs_pv_slice* ptrs = a.take<cubbs_pv_slice>(np);
  out->bbs_signature = a.copy(t.f[0]);
  out->bbs_header = a.copy(t.f[1]);
  out->public_key = a.copy(t.f[2]);
  out->hmac_key = a.copy(t.f[3]);
  for (size_t k = 0; k < np; k++) ptrs[k] = a.copy(t.pointers[k]);
  out->mandatory_pointers = ptrs;
  out->n_mandatory_pointers = np;
  return CUBBS_PV_OK;
}

static int32_t parse_derived_impl(const char* proof_value, size_t len, void* arena,
                                                  size_t arena_cap, size_t* arena_needed,
                                                  cubbs_pv_derived* out) {
  if (!arena_needed || !out || (!proof_value && len)) return CUBBS_PV_ERR_ARGUMENT;
  *arena_needed = 0;
  ParseScratch& t = tls;
  int32_t rc = unwrap(proof_value, len, PREFIX_DERIVED, t.bytes);
  if (rc) return rc;
  Reader r{t.bytes.data() + 3, t.bytes.size() - 3};
  Run& proof = t.f[0];
  Run& ph = t.f[1];
  const bool ok = r.tuplev(5, [&](int k) {
    switch (k) {
      case 0: return r.bytesv(proof);
      case 1: return r.u64_mapv(t.labels);
      case 2: return r.u64_listv(t.mandatory);
      case 3: return r.u64_listv(t.selective);
      default: return r.bytesv(ph);
    }
  });
  if (!ok) return CUBBS_PV_ERR_CBOR;
  if (!proof.size()) return CUBBS_PV_ERR_INVALID_DERIVED;  // lib.rs:155-160
  // BTreeMap<u64, u64>: ascending keys, the last of equal keys wins
  auto& l = t.labels;
  btree_insert_order(
      l, [](const auto& a, const auto& b) { return a.first < b.first; },
      [](const auto& a, const auto& b) { return a.first == b.first; });
  const size_t nl = l.size();
  // decompressed, ordered as BTreeMap<String, String>
  Label small[16];
  std::vector<Label> big;
  Label* lab = small;
  if (nl > 16) {
    big.resize(nl);
    lab = big.data();
  }
  for (size_t k = 0; k < nl; k++) {
    memcpy(lab[k].k, "c14n", 4);
    lab[k].kl = (uint8_t)(4 + decimal(lab[k].k + 4, l[k].first));
    lab[k].v[0] = 'b';
    lab[k].vl = (uint8_t)(1 + decimal(lab[k].v + 1, l[k].second));
  }
  std::sort(lab, lab + nl, label_less);
  size_t used = nl * sizeof(cubbs_pv_label);
  used += (t.mandatory.size() + t.selective.size()) * sizeof(uint64_t);
  used += proof.size() + ph.size();
  for (size_t k = 0; k < nl; k++) used += lab[k].kl + lab[k].vl;
  *arena_needed = arena_size(used);
  if (!arena || arena_cap < *arena_needed) return CUBBS_PV_ERR_INSUFFICIENT_OUTPUT;
  Arena a{aligned(arena)};
  cubbs_pv_label* labels = a.take<cubbs_pv_label>(nl);
  uint64_t* man = a.take<uint64_t>(t.mandatory.size());
  uint64_t* sel = a.take<uint64_t>(t.selective.size());
  if (!t.mandatory.empty()) memcpy(man, t.mandatory.data(), t.mandatory.size() * sizeof(uint64_t));
  if (!t.selective.empty()) memcpy(sel, t.selective.data(), t.selective.size() * sizeof(uint64_t));
  out->bbs_proof = a.copy(proof);
  for (size_t k = 0; k < nl; k++) {
    labels[k].key = a.copy((const uint8_t*)lab[k].k, lab[k].kl);
    labels[k].value = a.copy((const uint8_t*)lab[k].v, lab[k].vl);
  }
  out->label_map = labels;
  out->n_labels = nl;
  out->mandatory_indexes = man;
  out->n_mandatory_indexes = t.mandatory.size();
  out->selective_indexes = sel;
  out->n_selective_indexes = t.selective.size();
  out->presentation_header = a.copy(ph);
  return CUBBS_PV_OK;
}

// The C entries: no C++ exception crosses the boundary (the only one the
// codec can raise is an allocation failure).
template <class F>
static int32_t guarded(F&& f) {
  try {
    return f();
  } catch (...) {
    return CUBBS_PV_ERR_INTERNAL;
  }
}

extern "C" int32_t cubbs_proofvalue_serialize_base(const cubbs_pv_base* value, char* out,
                                                   size_t out_cap, size_t* out_len) {
  return guarded([&] { return serialize_base_impl(value, out, out_cap, out_len); });
}
extern "C" int32_t cubbs_proofvalue_serialize_derived(const cubbs_pv_derived* value, char* out,
                                                      size_t out_cap, size_t* out_len) {
  return guarded([&] { return serialize_derived_impl(value, out, out_cap, out_len); });
}
extern "C" int32_t cubbs_proofvalue_parse_base(const char* proof_value, size_t len, void* arena,
                                               size_t arena_cap, size_t* arena_needed, cubbs_pv_base* out) {
  return guarded([&] { return parse_base_impl(proof_value, len, arena, arena_cap, arena_needed, out); });
}
extern "C" int32_t cubbs_proofvalue_parse_derived(const char* proof_value, size_t len, void* arena,
                                                  size_t arena_cap, size_t* arena_needed,
                                                  cubbs_pv_derived* out) {
  return guarded([&] { return parse_derived_impl(proof_value, len, arena, arena_cap, arena_needed, out); });
}
