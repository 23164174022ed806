"""The bbs-2023 proofValue codec, over the C ABI of include/cubbs_proofvalue.h.

Mirrors the public API of crates/cubbs-bbs-proofvalue/src/lib.rs: the value
types BaseProofValue (lib.rs:37-44) and DerivedProofValue (lib.rs:46-54), the
four functions serialize_base / parse_base / serialize_derived /
parse_derived (lib.rs:73-121), and one error type whose `kind` names the
reference's Error variant (lib.rs:16-35).  All work happens in the native
library (cubbs-cred_amd/csrc/proofvalue.cpp); there is no Python fallback.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Dict, List

import cubbs_hip

OK, ERR_MULTIBASE, ERR_CBOR_PREFIX, ERR_INVALID_BASE, ERR_INVALID_DERIVED, ERR_BASE64, ERR_CBOR, \
    ERR_INSUFFICIENT_OUTPUT, ERR_ARGUMENT, ERR_INTERNAL = range(10)
KINDS = {ERR_MULTIBASE: "InvalidMultibasePrefix", ERR_CBOR_PREFIX: "InvalidCborPrefix",
         ERR_INVALID_BASE: "InvalidBaseProof", ERR_INVALID_DERIVED: "InvalidDerivedProof",
         ERR_BASE64: "Base64", ERR_CBOR: "Cbor", ERR_INSUFFICIENT_OUTPUT: "InsufficientOutput",
         ERR_ARGUMENT: "Argument", ERR_INTERNAL: "Internal"}
SYMBOLS = ["cubbs_proofvalue_serialize_base", "cubbs_proofvalue_parse_base",
           "cubbs_proofvalue_serialize_derived", "cubbs_proofvalue_parse_derived"]


class ProofValueError(Exception):
    def __init__(self, code):
        self.code = code
        self.kind = KINDS.get(code, "Unknown(%d)" % code)
        super().__init__(self.kind)


class Slice(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("len", ctypes.c_size_t)]


class CBase(ctypes.Structure):
    _fields_ = [("bbs_signature", Slice), ("bbs_header", Slice), ("public_key", Slice),
                ("hmac_key", Slice), ("mandatory_pointers", ctypes.POINTER(Slice)),
                ("n_mandatory_pointers", ctypes.c_size_t)]


class CLabel(ctypes.Structure):
    _fields_ = [("key", Slice), ("value", Slice)]


class CDerived(ctypes.Structure):
    _fields_ = [("bbs_proof", Slice), ("label_map", ctypes.POINTER(CLabel)), ("n_labels", ctypes.c_size_t),
                ("mandatory_indexes", ctypes.POINTER(ctypes.c_uint64)), ("n_mandatory_indexes", ctypes.c_size_t),
                ("selective_indexes", ctypes.POINTER(ctypes.c_uint64)), ("n_selective_indexes", ctypes.c_size_t),
                ("presentation_header", Slice)]


@dataclass
class BaseProofValue:
    bbs_signature: bytes
    bbs_header: bytes
    public_key: bytes
    hmac_key: bytes
    mandatory_pointers: List[str] = field(default_factory=list)


@dataclass
class DerivedProofValue:
    bbs_proof: bytes
    label_map: Dict[str, str] = field(default_factory=dict)
    mandatory_indexes: List[int] = field(default_factory=list)
    selective_indexes: List[int] = field(default_factory=list)
    presentation_header: bytes = b""


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = cubbs_hip.lib()
        for name in ("cubbs_proofvalue_serialize_base", "cubbs_proofvalue_serialize_derived"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_int32
            f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        for name in ("cubbs_proofvalue_parse_base", "cubbs_proofvalue_parse_derived"):
            f = getattr(_lib, name)
            f.restype = ctypes.c_int32
            f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                          ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
    return _lib


class _Keep:
    """Holds the ctypes buffers an input struct points at."""

    def __init__(self):
        self.refs = []

    def slice(self, b):
        b = bytes(b)
        buf = ctypes.create_string_buffer(b, len(b)) if b else None
        self.refs.append(buf)
        return Slice(ctypes.cast(buf, ctypes.c_void_p) if buf is not None else None, len(b))


def _text_bytes(s):
    return s.encode("utf-8", errors="surrogatepass") if isinstance(s, str) else bytes(s)


def _serialize(fn, cvalue):
    need = ctypes.c_size_t(0)
    rc = fn(ctypes.byref(cvalue), None, 0, ctypes.byref(need))
    if rc != ERR_INSUFFICIENT_OUTPUT:
        raise ProofValueError(rc)
    out = ctypes.create_string_buffer(need.value)
    rc = fn(ctypes.byref(cvalue), out, need.value, ctypes.byref(need))
    if rc != OK:
        raise ProofValueError(rc)
    return out.raw[:need.value].decode("ascii")


def serialize_base(value: BaseProofValue) -> str:
    """lib.rs:73-83."""
    k = _Keep()
    ptrs = (Slice * max(1, len(value.mandatory_pointers)))(
        *[k.slice(_text_bytes(p)) for p in value.mandatory_pointers])
    c = CBase(k.slice(value.bbs_signature), k.slice(value.bbs_header), k.slice(value.public_key),
              k.slice(value.hmac_key), ptrs, len(value.mandatory_pointers))
    return _serialize(lib().cubbs_proofvalue_serialize_base, c)


def serialize_derived(value: DerivedProofValue) -> str:
    """lib.rs:98-108."""
    k = _Keep()
    items = list(value.label_map.items())
    labels = (CLabel * max(1, len(items)))(
        *[CLabel(k.slice(_text_bytes(a)), k.slice(_text_bytes(b))) for a, b in items])
    man = (ctypes.c_uint64 * max(1, len(value.mandatory_indexes)))(*value.mandatory_indexes)
    sel = (ctypes.c_uint64 * max(1, len(value.selective_indexes)))(*value.selective_indexes)
    c = CDerived(k.slice(value.bbs_proof), labels, len(items), man, len(value.mandatory_indexes),
                 sel, len(value.selective_indexes), k.slice(value.presentation_header))
    return _serialize(lib().cubbs_proofvalue_serialize_derived, c)


def _parse(fn, proof_value, cvalue):
    raw = proof_value.encode("utf-8", errors="surrogatepass") if isinstance(proof_value, str) else bytes(proof_value)
    need = ctypes.c_size_t(0)
    rc = fn(raw, len(raw), None, 0, ctypes.byref(need), ctypes.byref(cvalue))
    if rc != ERR_INSUFFICIENT_OUTPUT:
        raise ProofValueError(rc)
    arena = ctypes.create_string_buffer(need.value)
    rc = fn(raw, len(raw), arena, need.value, ctypes.byref(need), ctypes.byref(cvalue))
    if rc != OK:
        raise ProofValueError(rc)
    return arena


def _get(s):
    return ctypes.string_at(s.ptr, s.len) if s.len else b""


def parse_base(proof_value) -> BaseProofValue:
    """lib.rs:85-96."""
    c = CBase()
    arena = _parse(lib().cubbs_proofvalue_parse_base, proof_value, c)
    value = BaseProofValue(_get(c.bbs_signature), _get(c.bbs_header), _get(c.public_key), _get(c.hmac_key),
                           [_get(c.mandatory_pointers[i]).decode() for i in range(c.n_mandatory_pointers)])
    del arena
    return value


def parse_derived(proof_value) -> DerivedProofValue:
    """lib.rs:110-121."""
    c = CDerived()
    arena = _parse(lib().cubbs_proofvalue_parse_derived, proof_value, c)
    labels = {}
    for i in range(c.n_labels):
        labels[_get(c.label_map[i].key).decode()] = _get(c.label_map[i].value).decode()
    value = DerivedProofValue(_get(c.bbs_proof), labels,
                              [c.mandatory_indexes[i] for i in range(c.n_mandatory_indexes)],
                              [c.selective_indexes[i] for i in range(c.n_selective_indexes)],
                              _get(c.presentation_header))
    del arena
    return value
