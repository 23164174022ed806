/* W3C Data Integrity `bbs-2023` proofValue codec: the wire format that wraps
 * a BBS signature (base proof) or a BBS proof (derived proof) for transport.
 *
 * Each entry replaces one public function of the reference's Rust crate
 * crates/cubbs-bbs-proofvalue/src/lib.rs:
 *   cubbs_proofvalue_serialize_base     <- serialize_base     lib.rs:73-83
 *   cubbs_proofvalue_parse_base         <- parse_base         lib.rs:85-96
 *   cubbs_proofvalue_serialize_derived  <- serialize_derived  lib.rs:98-108
 *   cubbs_proofvalue_parse_derived      <- parse_derived      lib.rs:110-121
 * with the same accepted inputs, outputs and error variants (lib.rs:16-35).
 *
 * Encoding: "u" + base64url-no-pad( 3-byte CBOR tag prefix || CBOR array ).
 *   base:    prefix d9 5d 02, array [bytes sig(80), bytes header(64),
 *            bytes public_key(96), bytes hmac_key(32), [text pointer...]]
 *   derived: prefix d9 5d 03, array [bytes proof(non-empty),
 *            {uint c14n-index: uint b-index}, [uint...], [uint...], bytes ph]
 * Label maps travel compressed: "c14n<k>" -> "b<v>" becomes k -> v
 * (lib.rs:165-189).
 *
 * Host code only; no call touches a GPU.  Every entry is thread-safe; the
 * only memory kept between calls is per-thread scratch capacity, whose
 * contents never carry over from one call to the next. */
#ifndef CUBBS_PROOFVALUE_H
#define CUBBS_PROOFVALUE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result codes.  1-6 are the reference's Error variants (lib.rs:16-35). */
#define CUBBS_PV_OK 0
#define CUBBS_PV_ERR_MULTIBASE_PREFIX 1   /* Error::InvalidMultibasePrefix */
#define CUBBS_PV_ERR_CBOR_PREFIX 2        /* Error::InvalidCborPrefix */
#define CUBBS_PV_ERR_INVALID_BASE 3       /* Error::InvalidBaseProof */
#define CUBBS_PV_ERR_INVALID_DERIVED 4    /* Error::InvalidDerivedProof */
#define CUBBS_PV_ERR_BASE64 5             /* Error::Base64 */
#define CUBBS_PV_ERR_CBOR 6               /* Error::Cbor */
/* Not in the reference (its outputs are owned Rust values): */
#define CUBBS_PV_ERR_INSUFFICIENT_OUTPUT 7 /* output buffer too small; the size
                                              needed is written back */
#define CUBBS_PV_ERR_ARGUMENT 8            /* NULL pointer with a non-zero
                                              length, or a text field that is
                                              not UTF-8 (a Rust String is) */
#define CUBBS_PV_ERR_INTERNAL 9            /* host memory exhausted */

/* A byte or UTF-8 text run.  ptr may be NULL when len is 0. */
typedef struct {
  const uint8_t* ptr;
  size_t len;
} cubbs_pv_slice;

/* BaseProofValue (lib.rs:37-44). */
typedef struct {
  cubbs_pv_slice bbs_signature;         /* 80 bytes */
  cubbs_pv_slice bbs_header;            /* 64 bytes */
  cubbs_pv_slice public_key;            /* 96 bytes */
  cubbs_pv_slice hmac_key;              /* 32 bytes */
  const cubbs_pv_slice* mandatory_pointers; /* UTF-8 JSON pointers, in order */
  size_t n_mandatory_pointers;
} cubbs_pv_base;

/* One label-map entry: canonical label ("c14n0") -> HMAC label ("b2"). */
typedef struct {
  cubbs_pv_slice key;
  cubbs_pv_slice value;
} cubbs_pv_label;

/* DerivedProofValue (lib.rs:46-54).  On input the label map is read as the
 * reference's BTreeMap<String, String> would hold it: entries ordered by key
 * bytes, and of equal keys the last one given wins.  On output the entries
 * come in that order. */
typedef struct {
  cubbs_pv_slice bbs_proof;             /* non-empty */
  const cubbs_pv_label* label_map;
  size_t n_labels;
  const uint64_t* mandatory_indexes;
  size_t n_mandatory_indexes;
  const uint64_t* selective_indexes;
  size_t n_selective_indexes;
  cubbs_pv_slice presentation_header;
} cubbs_pv_derived;

/* Serialization: writes the proofValue string (ASCII, no terminating NUL)
 * to out and its length to *out_len.  When out_cap is too small nothing is
 * written to out, *out_len receives the length needed and the call returns
 * CUBBS_PV_ERR_INSUFFICIENT_OUTPUT.  Validation errors come first. */
int32_t cubbs_proofvalue_serialize_base(const cubbs_pv_base* value, char* out,
                                        size_t out_cap, size_t* out_len);
int32_t cubbs_proofvalue_serialize_derived(const cubbs_pv_derived* value,
                                           char* out, size_t out_cap,
                                           size_t* out_len);

/* Parsing: decodes proof_value[0, len) into *out, whose slices and arrays
 * point into the caller's arena (any alignment; the library aligns inside
 * it).  *arena_needed always receives the arena size this value needs once
 * it has decoded; when arena_cap is smaller the call returns
 * CUBBS_PV_ERR_INSUFFICIENT_OUTPUT and leaves *out untouched.  Decoding and
 * validation errors come first (and leave *arena_needed 0). */
int32_t cubbs_proofvalue_parse_base(const char* proof_value, size_t len,
                                    void* arena, size_t arena_cap,
                                    size_t* arena_needed, cubbs_pv_base* out);
int32_t cubbs_proofvalue_parse_derived(const char* proof_value, size_t len,
                                       void* arena, size_t arena_cap,
                                       size_t* arena_needed,
                                       cubbs_pv_derived* out);

#ifdef __cplusplus
}
#endif

#endif /* CUBBS_PROOFVALUE_H */
