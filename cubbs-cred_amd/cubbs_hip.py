"""ctypes binding of libcubbs_bbs_hip.so (the C ABI v2 of cuBBS-Cred plus the
extensions in include/cubbs_hip_ext.h).

This is the Python-side mirror of the reference's FFI layer
(crates/cubbs-bbs/src/gpu_abi.rs for the struct layouts and
crates/cubbs-bbs/src/gpu_runtime.rs for how batches are marshalled): the same
items, the same result codes, the same per-item error behaviour.  Loading
fails loudly when the shared library is missing; there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CUBBS_BBS_GPU_LIBRARY") or os.path.join(HERE, "libcubbs_bbs_hip.so")

ABI_VERSION = 2
SUCCESS, ERR_INVALID_INPUT, ERR_INSUFFICIENT_OUTPUT, ERR_INTERNAL = 0, 1, 2, 3
BATCH_VERIFY, BATCH_PROOF_VERIFY = 0, 1
BATCH_G1_MSM, BATCH_G1_MSM_COMMON_POINTS = 16, 17
# a Verify and a ProofVerify batch staged as one (cubbs_hip_batch_prepare_verify_set);
# a tag of this binding, not a kind value of the C ABI
BATCH_VERIFY_SET = "verify_set"
BATCH_SINGLE_STREAM, BATCH_AGGREGATE = 1, 2

API_ID = b"BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_"
MAP_DST = API_ID + b"MAP_MSG_TO_SCALAR_AS_HASH_"
H2S_DST = API_ID + b"H2S_"

# Every exported symbol declared in include/cubbs_bbs_gpu_abi.h and
# include/cubbs_hip_ext.h (checked by the CPU test suite).
ABI_SYMBOLS = [
    "cubbs_bbs_gpu_backend_info_v2",
    "cubbs_bbs_gpu_sign_batch_v2",
    "cubbs_bbs_gpu_verify_batch_v2",
    "cubbs_bbs_gpu_proof_gen_batch_v2",
    "cubbs_bbs_gpu_proof_verify_batch_v2",
    "cubbs_bbs_gpu_messages_to_scalar_batch_v2",
    "cubbs_bbs_gpu_sha256_batch_for_test_v2",
    "cubbs_bbs_gpu_expand_xmd_sha256_48_batch_for_test_v2",
    "cubbs_bbs_gpu_hash_to_scalar_sha256_batch_for_test_v2",
    "cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2",
    "cubbs_bbs_gpu_prepare_b_batch_v2",
    "cubbs_bbs_gpu_g1_msm_batch_v2",
    "cubbs_bbs_gpu_g1_msm_common_points_batch_v2",
    "cubbs_bbs_gpu_g1_msm_fused_common_points_batch_v2",
    "cubbs_bbs_gpu_g1_msm_common_points_plus_variable_batch_v2",
    "cubbs_bbs_gpu_proof_verify_msm_pair_batch_v2",
    "cubbs_bbs_gpu_g1_dual_msm_same_points_batch_v2",
    "cubbs_bbs_gpu_g1_dual_msm_same_points_lanes_batch_v2",
    "cubbs_bbs_gpu_g1_add_batch_v2",
    "cubbs_bbs_gpu_pairing_eq_batch_v2",
    "cubbs_bbs_gpu_pairing_eq_aggregate_v2",
    "cubbs_bbs_gpu_proof_gen_t2_prepare_batch_v2",
]
EXT_SYMBOLS = [
    "cubbs_hip_batch_prepare",
    "cubbs_hip_batch_prepare_flags",
    "cubbs_hip_batch_launch",
    "cubbs_hip_batch_sync",
    "cubbs_hip_batch_results",
    "cubbs_hip_batch_free",
    "cubbs_hip_batch_outputs",
    "cubbs_hip_batch_use_graph",
    "cubbs_hip_server_timing",
    "cubbs_hip_autotune",
    "cubbs_hip_tune_load",
    "cubbs_hip_batch_profile",
    "cubbs_hip_batch_shape",
    "cubbs_hip_batch_group_checks",
    "cubbs_hip_batch_prepare_verify_set",
    "cubbs_hip_test_g1_mul",
    "cubbs_hip_test_pairing_gt",
    "cubbs_hip_test_g2_roundtrip",
    "cubbs_hip_generators",
    "cubbs_hip_sk_to_pk",
    "cubbs_hip_fp_mul_bench",
    "cubbs_hip_mad_peak_bench",
    "cubbs_hip_proof_gen_seeded_batch",
    "cubbs_hip_proof_gen_with_scalars_batch",
    "cubbs_hip_server_create",
    "cubbs_hip_server_submit",
    "cubbs_hip_server_wait",
    "cubbs_hip_server_run_open_loop",
    "cubbs_hip_server_stats",
    "cubbs_hip_server_queue_stats",
    "cubbs_hip_server_destroy",
]


class ByteSlice(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_size_t)]


class ByteBuffer(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("capacity", ctypes.c_size_t), ("len", ctypes.c_size_t)]


class IndexSlice(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_size_t)]


class Result(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int32), ("value", ctypes.c_uint8), ("bytes_written", ctypes.c_size_t)]


class BackendInfo(ctypes.Structure):
    _fields_ = [("backend_kind", ctypes.c_uint32), ("hardware_acceleration", ctypes.c_uint8)]


class SignItem(ctypes.Structure):
    _fields_ = [("secret_key", ByteSlice), ("public_key", ByteSlice), ("header", ByteSlice),
                ("messages", ctypes.c_void_p), ("message_count", ctypes.c_size_t),
                ("signature_out", ByteBuffer)]


class VerifyItem(ctypes.Structure):
    _fields_ = [("public_key", ByteSlice), ("signature", ByteSlice), ("header", ByteSlice),
                ("messages", ctypes.c_void_p), ("message_count", ctypes.c_size_t)]


class ProofGenItem(ctypes.Structure):
    _fields_ = [("public_key", ByteSlice), ("signature", ByteSlice), ("header", ByteSlice),
                ("presentation_header", ByteSlice), ("messages", ctypes.c_void_p),
                ("message_count", ctypes.c_size_t), ("disclosed_indexes", IndexSlice),
                ("proof_out", ByteBuffer)]


class ProofVerifyItem(ctypes.Structure):
    _fields_ = [("public_key", ByteSlice), ("proof", ByteSlice), ("header", ByteSlice),
                ("presentation_header", ByteSlice), ("disclosed_messages", ctypes.c_void_p),
                ("disclosed_count", ctypes.c_size_t), ("disclosed_indexes", IndexSlice),
                ("total_message_count", ctypes.c_size_t)]


class MessagesToScalarItem(ctypes.Structure):
    _fields_ = [("message", ByteSlice), ("dst", ByteSlice), ("scalar_out", ByteBuffer)]


class SignEItem(ctypes.Structure):
    _fields_ = [("secret_key", ByteSlice), ("domain", ByteSlice), ("messages", ctypes.c_void_p),
                ("message_scalar_offset", ctypes.c_size_t), ("message_count", ctypes.c_size_t),
                ("message_scalars_out", ByteBuffer), ("scalar_out", ByteBuffer)]


class PrepareBItem(ctypes.Structure):
    _fields_ = [("public_key", ByteSlice), ("header", ByteSlice), ("message_scalars", ctypes.c_void_p),
                ("message_count", ctypes.c_size_t), ("b_out", ByteBuffer)]


class G1MsmItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("scalars", ctypes.c_void_p), ("term_count", ctypes.c_size_t),
                ("point_out", ByteBuffer)]


class CommonPointsItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("term_count", ctypes.c_size_t),
                ("scalar_rows", ctypes.c_void_p), ("scalar_rows_be", ctypes.c_void_p),
                ("scalar_row_stride", ctypes.c_size_t), ("output_count", ctypes.c_size_t),
                ("point_outs", ctypes.c_void_p)]


class PlusVariableItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("term_count", ctypes.c_size_t),
                ("scalar_rows", ctypes.c_void_p), ("variable_points", ctypes.c_void_p),
                ("variable_scalars", ctypes.c_void_p), ("output_count", ctypes.c_size_t),
                ("point_outs", ctypes.c_void_p)]


class PvMsmPairItem(ctypes.Structure):
    _fields_ = [("common_points", ctypes.c_void_p), ("common_term_count", ctypes.c_size_t),
                ("common_scalar_rows", ctypes.c_void_p), ("proof_points", ctypes.c_void_p),
                ("proof_scalars", ctypes.c_void_p), ("output_count", ctypes.c_size_t),
                ("t1_outs", ctypes.c_void_p), ("t2_outs", ctypes.c_void_p)]


class DualMsmItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("lhs_scalars", ctypes.c_void_p),
                ("rhs_scalars", ctypes.c_void_p), ("term_count", ctypes.c_size_t),
                ("lhs_out", ByteBuffer), ("rhs_out", ByteBuffer)]


class DualMsmLanesItem(ctypes.Structure):
    _fields_ = [("points", ctypes.c_void_p), ("lhs_scalars", ctypes.c_void_p),
                ("rhs_scalars", ctypes.c_void_p), ("term_count", ctypes.c_size_t),
                ("lane_count", ctypes.c_size_t), ("lhs_outs", ByteBuffer), ("rhs_outs", ByteBuffer)]


class G1AddItem(ctypes.Structure):
    _fields_ = [("lhs", ByteSlice), ("rhs", ByteSlice), ("point_out", ByteBuffer)]


class PairingEqItem(ctypes.Structure):
    _fields_ = [("lhs_g1", ByteSlice), ("lhs_g2", ByteSlice), ("rhs_g1", ByteSlice), ("rhs_g2", ByteSlice)]


class PairingEqScaledItem(ctypes.Structure):
    _fields_ = [("lhs_g1", ByteSlice), ("lhs_g2", ByteSlice), ("rhs_g1", ByteSlice), ("rhs_g2", ByteSlice),
                ("coefficient", ByteSlice)]


class T2PrepareItem(ctypes.Structure):
    _fields_ = [("compressed_b", ByteSlice), ("signature_a", ByteSlice), ("scalars_be", ByteSlice * 7),
                ("seed", ByteSlice), ("d_out", ByteBuffer), ("abar_out", ByteBuffer),
                ("bbar_out", ByteBuffer), ("t1_out", ByteBuffer)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_uint32), ("max_batch", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("flags", ctypes.c_uint32),
                ("coalesce_us", ctypes.c_uint32), ("queue_capacity", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


SERVER_GRAPHS, SERVER_AGGREGATE = 1, 2


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("item", ctypes.c_void_p), ("seed32", ctypes.c_void_p)]


REQ_VERIFY, REQ_PROOF_VERIFY, REQ_SIGN, REQ_PROOF_GEN = 0, 1, 2, 3


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcubbs_bbs_hip.so not built (%s): run __graft_entry__.build()" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.cubbs_hip_batch_prepare.restype = ctypes.c_void_p
        _lib.cubbs_hip_batch_prepare.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_size_t, ctypes.POINTER(ctypes.c_int32)]
        _lib.cubbs_hip_batch_prepare_flags.restype = ctypes.c_void_p
        _lib.cubbs_hip_batch_prepare_verify_set.restype = ctypes.c_void_p
        _lib.cubbs_hip_batch_prepare_verify_set.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_size_t,
                                                            ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32,
                                                            ctypes.POINTER(ctypes.c_int32)]
        _lib.cubbs_hip_batch_prepare_flags.argtypes = [ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p,
                                                       ctypes.c_size_t, ctypes.c_uint32,
                                                       ctypes.POINTER(ctypes.c_int32)]
        for name in ("cubbs_hip_batch_launch", "cubbs_hip_batch_sync"):
            getattr(_lib, name).argtypes = [ctypes.c_void_p]
        _lib.cubbs_hip_batch_results.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.cubbs_hip_batch_free.argtypes = [ctypes.c_void_p]
        _lib.cubbs_hip_batch_free.restype = None
        _lib.cubbs_hip_batch_use_graph.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        _lib.cubbs_hip_batch_outputs.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                 ctypes.POINTER(ctypes.c_size_t)]
        _lib.cubbs_hip_batch_profile.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                                 ctypes.c_size_t, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_void_p]
        _lib.cubbs_hip_batch_shape.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        _lib.cubbs_hip_batch_group_checks.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 2
        _lib.cubbs_hip_batch_work.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.cubbs_hip_server_create.restype = ctypes.c_void_p
        _lib.cubbs_hip_server_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]
        _lib.cubbs_hip_server_submit.restype = ctypes.c_int64
        _lib.cubbs_hip_server_submit.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _lib.cubbs_hip_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        _lib.cubbs_hip_server_run_open_loop.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p,
                                                        ctypes.c_void_p, ctypes.c_void_p]
        _lib.cubbs_hip_server_stats.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        _lib.cubbs_hip_server_queue_stats.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [ctypes.c_size_t]
        _lib.cubbs_hip_server_destroy.argtypes = [ctypes.c_void_p]
        _lib.cubbs_hip_server_destroy.restype = None
    return _lib


class _Keep:
    """Holds ctypes buffers alive for the duration of a call."""

    def __init__(self):
        self.refs = []

    def slice(self, b):
        b = bytes(b)
        if not b:
            return ByteSlice(None, 0)
        buf = ctypes.create_string_buffer(b, len(b))
        self.refs.append(buf)
        return ByteSlice(ctypes.cast(buf, ctypes.c_void_p), len(b))

    def slices(self, items):
        arr = (ByteSlice * max(1, len(items)))(*[self.slice(m) for m in items])
        self.refs.append(arr)
        return ctypes.cast(arr, ctypes.c_void_p) if items else None

    def indexes(self, idx):
        arr = (ctypes.c_size_t * max(1, len(idx)))(*idx)
        self.refs.append(arr)
        return IndexSlice(ctypes.cast(arr, ctypes.c_void_p) if idx else None, len(idx))

    def out(self, cap):
        buf = ctypes.create_string_buffer(max(1, cap))
        self.refs.append(buf)
        return buf, ByteBuffer(ctypes.cast(buf, ctypes.c_void_p), cap, 0)


def backend_info():
    info = BackendInfo()
    rc = lib().cubbs_bbs_gpu_backend_info_v2(ABI_VERSION, ctypes.byref(info))
    return rc, info.backend_kind, info.hardware_acceleration


def build_verify_items(reqs, keep):
    """reqs: list of (pk, sig, header, [messages])."""
    items = (VerifyItem * max(1, len(reqs)))()
    for i, (pk, sig, hdr, msgs) in enumerate(reqs):
        items[i].public_key = keep.slice(pk)
        items[i].signature = keep.slice(sig)
        items[i].header = keep.slice(hdr)
        items[i].messages = keep.slices(msgs)
        items[i].message_count = len(msgs)
    keep.refs.append(items)
    return items


def verify_batch(reqs):
    """Returns (status, [(code, valid)])."""
    keep = _Keep()
    items = build_verify_items(reqs, keep)
    res = (Result * max(1, len(reqs)))()
    rc = lib().cubbs_bbs_gpu_verify_batch_v2(ABI_VERSION, items, ctypes.c_size_t(len(reqs)), res)
    return rc, [(res[i].code, bool(res[i].value)) for i in range(len(reqs))]


def build_proof_verify_items(reqs, keep):
    """reqs: list of (pk, proof, header, ph, disclosed_msgs, disclosed_idx, total)."""
    items = (ProofVerifyItem * max(1, len(reqs)))()
    for i, (pk, proof, hdr, ph, dmsgs, didx, total) in enumerate(reqs):
        items[i].public_key = keep.slice(pk)
        items[i].proof = keep.slice(proof)
        items[i].header = keep.slice(hdr)
        items[i].presentation_header = keep.slice(ph)
        items[i].disclosed_messages = keep.slices(dmsgs)
        items[i].disclosed_count = len(dmsgs)
        items[i].disclosed_indexes = keep.indexes(didx)
        items[i].total_message_count = total
    keep.refs.append(items)
    return items


def proof_verify_batch(reqs):
    """Returns (status, [(code, valid)])."""
    keep = _Keep()
    items = build_proof_verify_items(reqs, keep)
    res = (Result * max(1, len(reqs)))()
    rc = lib().cubbs_bbs_gpu_proof_verify_batch_v2(ABI_VERSION, items, ctypes.c_size_t(len(reqs)), res)
    return rc, [(res[i].code, bool(res[i].value)) for i in range(len(reqs))]


def sign_batch(reqs):
    """reqs: list of (sk, pk, header, [messages]) -> (status, [(code, sig or None)])."""
    keep = _Keep()
    items = (SignItem * max(1, len(reqs)))()
    outs = []
    for i, (sk, pk, hdr, msgs) in enumerate(reqs):
        items[i].secret_key = keep.slice(sk)
        items[i].public_key = keep.slice(pk)
        items[i].header = keep.slice(hdr)
        items[i].messages = keep.slices(msgs)
        items[i].message_count = len(msgs)
        buf, bb = keep.out(80)
        items[i].signature_out = bb
        outs.append(buf)
    res = (Result * max(1, len(reqs)))()
    rc = lib().cubbs_bbs_gpu_sign_batch_v2(ABI_VERSION, items, ctypes.c_size_t(len(reqs)), res)
    return rc, [(res[i].code, outs[i].raw[:items[i].signature_out.len] if res[i].code == 0 else None)
                for i in range(len(reqs))]


def proof_gen_batch(reqs, seeds=None, scalars=None):
    """reqs: list of (pk, sig, header, ph, [messages], disclosed_idx).

    seeds: list of 32-byte seeds (ChaCha20 blinding) or None; scalars: list of
    lists of 32-byte blinding scalars (explicit) or None (OS randomness).
    Returns (status, [(code, proof or None)])."""
    keep = _Keep()
    items = (ProofGenItem * max(1, len(reqs)))()
    outs = []
    for i, (pk, sig, hdr, ph, msgs, didx) in enumerate(reqs):
        items[i].public_key = keep.slice(pk)
        items[i].signature = keep.slice(sig)
        items[i].header = keep.slice(hdr)
        items[i].presentation_header = keep.slice(ph)
        items[i].messages = keep.slices(msgs)
        items[i].message_count = len(msgs)
        items[i].disclosed_indexes = keep.indexes(didx)
        cap = 272 + 32 * max(0, len(msgs) - len(didx))
        buf, bb = keep.out(cap)
        items[i].proof_out = bb
        outs.append(buf)
    res = (Result * max(1, len(reqs)))()
    n = ctypes.c_size_t(len(reqs))
    if scalars is not None:
        flat = b"".join(b"".join(x) for x in scalars)
        offs, cur = [], 0
        for x in scalars:
            offs.append(cur)
            cur += len(x)
        offarr = (ctypes.c_size_t * max(1, len(offs)))(*offs)
        sbuf = ctypes.create_string_buffer(flat, max(1, len(flat)))
        rc = lib().cubbs_hip_proof_gen_with_scalars_batch(ABI_VERSION, items, n, sbuf, offarr, res)
    elif seeds is not None:
        sbuf = ctypes.create_string_buffer(b"".join(seeds), max(1, 32 * len(seeds)))
        rc = lib().cubbs_hip_proof_gen_seeded_batch(ABI_VERSION, items, n, sbuf, res)
    else:
        rc = lib().cubbs_bbs_gpu_proof_gen_batch_v2(ABI_VERSION, items, n, res)
    return rc, [(res[i].code, outs[i].raw[:items[i].proof_out.len] if res[i].code == 0 else None)
                for i in range(len(reqs))]


def sk_to_pk(sks):
    n = len(sks)
    out = ctypes.create_string_buffer(96 * max(1, n))
    ok = ctypes.create_string_buffer(max(1, n))
    rc = lib().cubbs_hip_sk_to_pk(b"".join(sks), ctypes.c_size_t(n), out, ok)
    return rc, [(out.raw[96 * i:96 * i + 96] if ok.raw[i] else None) for i in range(n)]


def keygen(ikms, info=b""):
    """SK = hash_to_scalar(IKM || I2OSP(len(info), 2) || info, api_id || "KEYGEN_DST_"),
    PK = SK * G2 (crates/cubbs-bbs-protocol/src/bbsplus/keys.rs:214-260), on the device."""
    inputs = [ikm + len(info).to_bytes(2, "big") + info for ikm in ikms]
    rc, sks = hash_to_scalar_batch(inputs, API_ID + b"KEYGEN_DST_")
    if rc:
        return rc, []
    rc, pks = sk_to_pk(sks)
    return rc, list(zip(sks, pks))


def messages_to_scalars(msgs, dst=MAP_DST):
    keep = _Keep()
    items = (MessagesToScalarItem * max(1, len(msgs)))()
    outs = []
    for i, m in enumerate(msgs):
        items[i].message = keep.slice(m)
        items[i].dst = keep.slice(dst)
        buf, bb = keep.out(32)
        items[i].scalar_out = bb
        outs.append(buf)
    res = (Result * max(1, len(msgs)))()
    rc = lib().cubbs_bbs_gpu_messages_to_scalar_batch_v2(ABI_VERSION, items, ctypes.c_size_t(len(msgs)), res)
    return rc, [o.raw[:32] for o in outs]


def messages_to_scalars_noalloc(msgs, dst=MAP_DST):
    """cubbs_hip_messages_to_scalar_noalloc_batch: (status, scalars, results)."""
    keep = _Keep()
    items = (MessagesToScalarItem * max(1, len(msgs)))()
    outs = []
    for i, m in enumerate(msgs):
        items[i].message = keep.slice(m)
        items[i].dst = keep.slice(dst)
        buf, bb = keep.out(32)
        items[i].scalar_out = bb
        outs.append(buf)
    res = (Result * max(1, len(msgs)))()
    rc = lib().cubbs_hip_messages_to_scalar_noalloc_batch(ABI_VERSION, items, ctypes.c_size_t(len(msgs)), res)
    return rc, [o.raw[:32] for o in outs], [(res[i].code, res[i].bytes_written) for i in range(len(msgs))]


def _packed(msgs):
    data = b"".join(msgs)
    offs, cur = [], 0
    for m in msgs:
        offs.append(cur)
        cur += len(m)
    n = max(1, len(msgs))
    return (ctypes.create_string_buffer(data, max(1, len(data))),
            (ctypes.c_size_t * n)(*offs), (ctypes.c_size_t * n)(*[len(m) for m in msgs]))


def sha256_batch(msgs):
    data, offs, lens = _packed(msgs)
    out = ctypes.create_string_buffer(32 * max(1, len(msgs)))
    rc = lib().cubbs_bbs_gpu_sha256_batch_for_test_v2(ABI_VERSION, data, offs, lens, ctypes.c_size_t(len(msgs)), out)
    return rc, [out.raw[32 * i:32 * i + 32] for i in range(len(msgs))]


def expand_xmd48_batch(msgs, dst):
    data, offs, lens = _packed(msgs)
    out = ctypes.create_string_buffer(48 * max(1, len(msgs)))
    d = ctypes.create_string_buffer(dst, max(1, len(dst)))
    rc = lib().cubbs_bbs_gpu_expand_xmd_sha256_48_batch_for_test_v2(
        ABI_VERSION, data, offs, lens, ctypes.c_size_t(len(msgs)), d, ctypes.c_size_t(len(dst)), out)
    return rc, [out.raw[48 * i:48 * i + 48] for i in range(len(msgs))]


def hash_to_scalar_batch(msgs, dst):
    data, offs, lens = _packed(msgs)
    out = ctypes.create_string_buffer(32 * max(1, len(msgs)))
    d = ctypes.create_string_buffer(dst, max(1, len(dst)))
    rc = lib().cubbs_bbs_gpu_hash_to_scalar_sha256_batch_for_test_v2(
        ABI_VERSION, data, offs, lens, ctypes.c_size_t(len(msgs)), d, ctypes.c_size_t(len(dst)), out)
    return rc, [out.raw[32 * i:32 * i + 32] for i in range(len(msgs))]


def g1_mul(points, scalars):
    n = len(points)
    out = ctypes.create_string_buffer(48 * max(1, n))
    ok = ctypes.create_string_buffer(max(1, n))
    rc = lib().cubbs_hip_test_g1_mul(b"".join(points), b"".join(scalars), ctypes.c_size_t(n), out, ok)
    return rc, [(out.raw[48 * i:48 * i + 48] if ok.raw[i] else None) for i in range(n)]


def pairing_gt(g1s, g2s):
    n = len(g1s)
    out = ctypes.create_string_buffer(576 * max(1, n))
    ok = ctypes.create_string_buffer(max(1, n))
    rc = lib().cubbs_hip_test_pairing_gt(b"".join(g1s), b"".join(g2s), ctypes.c_size_t(n), out, ok)
    return rc, [(out.raw[576 * i:576 * i + 576] if ok.raw[i] else None) for i in range(n)]


def g2_roundtrip(g2s):
    n = len(g2s)
    out = ctypes.create_string_buffer(96 * max(1, n))
    ok = ctypes.create_string_buffer(max(1, n))
    rc = lib().cubbs_hip_test_g2_roundtrip(b"".join(g2s), ctypes.c_size_t(n), out, ok)
    return rc, [(out.raw[96 * i:96 * i + 96] if ok.raw[i] else None) for i in range(n)]


def generators(count):
    out = ctypes.create_string_buffer(48 * count)
    rc = lib().cubbs_hip_generators(ctypes.c_size_t(count), out)
    return rc, [out.raw[48 * i:48 * i + 48] for i in range(count)]


def autotune(max_log2_terms=16):
    """Measured launch-shape tuning; returns (status, report text)."""
    buf = ctypes.create_string_buffer(8192)
    rc = lib().cubbs_hip_autotune(ctypes.c_uint32(max_log2_terms), buf, ctypes.c_size_t(8192))
    return rc, buf.value.decode()


def tune_load(text):
    return lib().cubbs_hip_tune_load(text.encode())


RESIDENT_FIELDS = ("resident_bytes", "pk_capacity", "pk_entries", "pk_pinned", "generator_slots",
                   "table_slots", "table_max_slots", "wide_slots", "retired_buffers", "live_batches")


def resident_stats():
    """The library's resident device state (cubbs_hip_resident_stats)."""
    v = (ctypes.c_uint64 * len(RESIDENT_FIELDS))()
    rc = lib().cubbs_hip_resident_stats(v)
    if rc:
        raise RuntimeError("resident_stats failed %d" % rc)
    return dict(zip(RESIDENT_FIELDS, list(v)))


def force_shapes(pip_c=0, pip_s=0, comb_chunk=0):
    """Launch shapes forced on the calling thread for its next calls (0 = not forced)."""
    return lib().cubbs_hip_test_force_shapes(ctypes.c_uint32(pip_c), ctypes.c_uint32(pip_s),
                                             ctypes.c_uint32(comb_chunk))


def fp_mul_bench(lanes, iters):
    ms = ctypes.c_double()
    rc = lib().cubbs_hip_fp_mul_bench(ctypes.c_uint32(lanes), ctypes.c_uint32(iters), ctypes.byref(ms))
    return rc, ms.value


def mad_peak_bench(lanes, iters):
    """Elapsed ms of lanes * iters * 64 independent v_mad_u64_u32."""
    ms = ctypes.c_double()
    rc = lib().cubbs_hip_mad_peak_bench(ctypes.c_uint32(lanes), ctypes.c_uint32(iters), ctypes.byref(ms))
    return rc, ms.value


class StagedBatch:
    """A batch whose inputs stay resident on the device (prepare once, launch many)."""

    def __init__(self, kind, reqs, single_stream=False, aggregate=False, comb_chunk=0, supergroup=0,
                 straus=0):
        """aggregate: False, True (the tuned group size, else 8) or a group
        size 2..8; supergroup: groups per supergroup check (2..8, with
        aggregate; 0 = the tuned choice when the group size is left open,
        else none); comb_chunk: fixed-base comb terms per lane (Verify /
        ProofVerify; 0 = the tuned choice, else 4); straus: signatures per
        product job of an aggregated Verify group's sums (1..4, with
        aggregate; 0 = two products per item).  kind BATCH_VERIFY_SET:
        reqs = (verify requests, proof-verify requests), staged as one batch
        whose results are the Verify results followed by the others."""
        self.keep = _Keep()
        flags = ((BATCH_SINGLE_STREAM if single_stream else 0) |
                 (BATCH_AGGREGATE if aggregate else 0) |
                 (0 if aggregate is True or not aggregate else (int(aggregate) & 15) << 8) |
                 ((int(comb_chunk) & 127) << 12) |
                 ((int(supergroup) & 15) << 20) |
                 ((int(straus) & 15) << 24))
        st = ctypes.c_int32(0)
        self.kind = kind
        if kind == BATCH_VERIFY_SET:
            vreqs, preqs = reqs
            vitems = build_verify_items(vreqs, self.keep)
            pitems = build_proof_verify_items(preqs, self.keep)
            self.n = len(vreqs) + len(preqs)
            self.h = lib().cubbs_hip_batch_prepare_verify_set(
                ABI_VERSION, ctypes.cast(vitems, ctypes.c_void_p), ctypes.c_size_t(len(vreqs)),
                ctypes.cast(pitems, ctypes.c_void_p), ctypes.c_size_t(len(preqs)), flags, ctypes.byref(st))
            self.keep = None
            if not self.h:
                raise RuntimeError("batch prepare failed with status %d" % st.value)
            return
        if kind == BATCH_VERIFY:
            items = build_verify_items(reqs, self.keep)
        elif kind == BATCH_PROOF_VERIFY:
            items = build_proof_verify_items(reqs, self.keep)
        elif kind == BATCH_G1_MSM:
            items, _ = _msm_items(reqs, self.keep)
        elif kind == BATCH_G1_MSM_COMMON_POINTS:
            items, _ = _common_items(reqs, self.keep, packed=True, outputs=False)
        else:
            raise ValueError("unsupported batch kind")
        # results: one per item, or one per output row of the common-point groups
        self.n = sum(len(rows) for _, rows in reqs) if kind == BATCH_G1_MSM_COMMON_POINTS else len(reqs)
        self.h = lib().cubbs_hip_batch_prepare_flags(ABI_VERSION, kind, ctypes.cast(items, ctypes.c_void_p),
                                                     ctypes.c_size_t(len(reqs)), flags, ctypes.byref(st))
        self.keep = None  # items are only read during prepare
        if not self.h:
            raise RuntimeError("batch prepare failed with status %d" % st.value)

    def launch(self):
        rc = lib().cubbs_hip_batch_launch(self.h)
        if rc:
            raise RuntimeError("launch failed %d" % rc)

    def sync(self):
        rc = lib().cubbs_hip_batch_sync(self.h)
        if rc:
            raise RuntimeError("sync failed %d" % rc)

    def results(self):
        res = (Result * max(1, self.n))()
        rc = lib().cubbs_hip_batch_results(self.h, res)
        if rc:
            raise RuntimeError("results failed %d" % rc)
        if self.kind in (BATCH_G1_MSM, BATCH_G1_MSM_COMMON_POINTS):
            return [(res[i].code, res[i].value) for i in range(self.n)]
        return [(res[i].code, bool(res[i].value)) for i in range(self.n)]

    def results_raw(self):
        """The result array as bytes (cheap to compare between launches)."""
        res = (Result * max(1, self.n))()
        rc = lib().cubbs_hip_batch_results(self.h, res)
        if rc:
            raise RuntimeError("results failed %d" % rc)
        return bytes(res)

    def use_graph(self, on=True):
        """Capture the launch sequence into a HIP graph (on) or drop it (off)."""
        rc = lib().cubbs_hip_batch_use_graph(self.h, 1 if on else 0)
        if rc:
            raise RuntimeError("graph capture failed %d" % rc)

    def outputs(self):
        """MSM kinds, after results(): the 48-byte output points in result order."""
        buf = ctypes.create_string_buffer(48 * max(1, self.n))
        ln = ctypes.c_size_t(0)
        rc = lib().cubbs_hip_batch_outputs(self.h, buf, ctypes.c_size_t(48 * self.n), ctypes.byref(ln))
        if rc:
            raise RuntimeError("outputs failed %d" % rc)
        return [buf.raw[48 * i:48 * i + 48] for i in range(ln.value // 48)]

    def profile(self, iters):
        names = ctypes.create_string_buffer(4096)
        ms = (ctypes.c_double * 64)()
        total = ctypes.c_double()
        k = lib().cubbs_hip_batch_profile(self.h, iters, names, 4096, ms, 64, ctypes.byref(total))
        if k < 0:
            raise RuntimeError("profile failed %d" % k)
        nm = names.value.decode().split("\n")[:k]
        return total.value, dict(zip(nm, [ms[i] for i in range(k)]))

    def shape(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        lib().cubbs_hip_batch_shape(self.h, *[ctypes.byref(x) for x in v])
        return dict(zip(["items", "msm_terms", "comb_madds", "pairings"], [x.value for x in v]))

    WORK_FIELDS = ("items", "comb_terms", "comb_madds", "glv_products", "agg_products", "decoded_points",
                   "point_adds", "pair_checks", "group_checks", "pair_checks_run", "sha_blocks",
                   "straus_sums", "straus_terms", "straus_short_sums", "straus_short_terms")

    def work(self):
        """Algorithmic work units of the batch (cubbs_hip_batch_work); pair_checks_run
        reflects the last results()."""
        v = (ctypes.c_uint64 * len(self.WORK_FIELDS))()
        rc = lib().cubbs_hip_batch_work(self.h, v)
        if rc:
            raise RuntimeError("work failed %d" % rc)
        return dict(zip(self.WORK_FIELDS, list(v)))

    def group_checks(self):
        """(aggregated group checks in the batch, how many held at the last
        results fetch)."""
        v = [ctypes.c_uint64() for _ in range(2)]
        lib().cubbs_hip_batch_group_checks(self.h, *[ctypes.byref(x) for x in v])
        return v[0].value, v[1].value

    def free(self):
        if self.h:
            lib().cubbs_hip_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _request_item(kind, payload, keep):
    """One ABI item for a server request (output buffers are not needed)."""
    if kind == REQ_VERIFY:
        return build_verify_items([payload], keep)
    if kind == REQ_PROOF_VERIFY:
        return build_proof_verify_items([payload], keep)
    if kind == REQ_SIGN:
        sk, pk, hdr, msgs = payload
        it = (SignItem * 1)()
        it[0].secret_key = keep.slice(sk)
        it[0].public_key = keep.slice(pk)
        it[0].header = keep.slice(hdr)
        it[0].messages = keep.slices(msgs)
        it[0].message_count = len(msgs)
        keep.refs.append(it)
        return it
    if kind == REQ_PROOF_GEN:
        pk, sig, hdr, ph, msgs, didx = payload
        it = (ProofGenItem * 1)()
        it[0].public_key = keep.slice(pk)
        it[0].signature = keep.slice(sig)
        it[0].header = keep.slice(hdr)
        it[0].presentation_header = keep.slice(ph)
        it[0].messages = keep.slices(msgs)
        it[0].message_count = len(msgs)
        it[0].disclosed_indexes = keep.indexes(didx)
        keep.refs.append(it)
        return it
    raise ValueError("unknown request kind %r" % kind)


def _out_cap(kind, payload):
    if kind == REQ_SIGN:
        return 80
    if kind == REQ_PROOF_GEN:
        return 272 + 32 * max(0, len(payload[4]) - len(payload[5]))
    return 0


class Server:
    """The library's serving loop (continuous batching over device slots).

    Requests are (kind, payload, seed) with payloads shaped as for
    verify_batch / proof_verify_batch / sign_batch / proof_gen_batch and seed a
    32-byte ProofGen blinding seed or None."""

    def __init__(self, slots=0, max_batch=0, hbm_budget=0, graphs=False, aggregate=False, coalesce_us=0,
                 queue_capacity=0):
        """coalesce_us: the window a partial batch waits to fill; queue_capacity:
        pending requests at most (submit blocks while full); 0 = none."""
        cfg = ServerConfig(slots, max_batch, hbm_budget,
                           (SERVER_GRAPHS if graphs else 0) | (SERVER_AGGREGATE if aggregate else 0),
                           coalesce_us, queue_capacity, 0)
        st = ctypes.c_int32(0)
        self.h = lib().cubbs_hip_server_create(ctypes.byref(cfg), ctypes.byref(st))
        if not self.h:
            raise RuntimeError("server create failed with status %d" % st.value)

    def _requests(self, pool, keep):
        reqs = (Request * max(1, len(pool)))()
        for i, (kind, payload, seed) in enumerate(pool):
            it = _request_item(kind, payload, keep)
            sd = None
            if seed is not None:
                buf = ctypes.create_string_buffer(bytes(seed), 32)
                keep.refs.append(buf)
                sd = ctypes.cast(buf, ctypes.c_void_p)
            reqs[i] = Request(kind, ctypes.cast(it, ctypes.c_void_p), sd)
        keep.refs.append(reqs)
        return reqs

    def submit(self, kind, payload, seed=None):
        keep = _Keep()
        reqs = self._requests([(kind, payload, seed)], keep)
        t = lib().cubbs_hip_server_submit(self.h, ctypes.byref(reqs[0]))
        if t < 0:
            raise RuntimeError("submit rejected with code %d" % -t)
        return t, _out_cap(kind, payload)

    def wait(self, ticket):
        """ticket as returned by submit -> (code, value, output bytes, latency_ns)."""
        t, cap = ticket
        res = Result()
        out = ctypes.create_string_buffer(max(1, cap))
        lat = ctypes.c_uint64()
        rc = lib().cubbs_hip_server_wait(self.h, t, ctypes.byref(res), out, cap, ctypes.byref(lat))
        if rc:
            raise RuntimeError("wait failed %d" % rc)
        return res.code, res.value, out.raw[:res.bytes_written], lat.value

    def run_open_loop(self, pool, order, arrivals_ns, outputs=False):
        """Submit pool[order[i]] at arrivals_ns[i]; returns (status, results,
        latencies_ns, outputs) with results [(code, value)]."""
        keep = _Keep()
        preqs = self._requests(pool, keep)
        n = len(order)
        reqs = (Request * max(1, n))()
        for i, k in enumerate(order):
            reqs[i] = preqs[k]
        arr = (ctypes.c_uint64 * max(1, n))(*arrivals_ns)
        res = (Result * max(1, n))()
        lat = (ctypes.c_uint64 * max(1, n))()
        outs_p = caps_p = None
        bufs = []
        if outputs:
            outs_p = (ctypes.c_void_p * max(1, n))()
            caps_p = (ctypes.c_size_t * max(1, n))()
            for i, k in enumerate(order):
                cap = _out_cap(pool[k][0], pool[k][1])
                b = ctypes.create_string_buffer(max(1, cap))
                bufs.append(b)
                outs_p[i] = ctypes.cast(b, ctypes.c_void_p)
                caps_p[i] = cap
        rc = lib().cubbs_hip_server_run_open_loop(self.h, reqs, arr, ctypes.c_size_t(n), res, lat,
                                                  outs_p, caps_p)
        results = [(res[i].code, res[i].value) for i in range(n)]
        outv = [bufs[i].raw[:res[i].bytes_written] for i in range(n)] if outputs else None
        return rc, results, [lat[i] for i in range(n)], outv

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        rc = lib().cubbs_hip_server_stats(self.h, *[ctypes.byref(x) for x in v])
        d = dict(zip(["batches", "requests", "max_concurrent_batches", "peak_slot_bytes"],
                     [x.value for x in v]))
        d["fault"] = rc
        pm, bm = ctypes.c_double(), ctypes.c_double()
        lib().cubbs_hip_server_timing.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib().cubbs_hip_server_timing(self.h, ctypes.byref(pm), ctypes.byref(bm))
        d["host_plan_ms_per_batch"] = round(pm.value, 3)
        d["batch_ms_per_batch"] = round(bm.value, 3)
        return d

    def queue_stats(self, hist_len=0):
        """Batcher statistics: requests queued now, summed queue wait (ns) of the
        dispatched requests, the largest and the last batch, and (hist_len > 0)
        batches per size, the last bucket holding every larger batch."""
        v = [ctypes.c_uint64() for _ in range(4)]
        hist = (ctypes.c_uint64 * max(1, hist_len))()
        rc = lib().cubbs_hip_server_queue_stats(self.h, *[ctypes.byref(x) for x in v],
                                                hist if hist_len else None, ctypes.c_size_t(hist_len))
        if rc < 0 or rc == ERR_INVALID_INPUT:
            raise RuntimeError("queue_stats failed %d" % rc)
        d = dict(zip(["queued", "total_wait_ns", "max_batch", "last_batch"], [x.value for x in v]))
        if hist_len:
            d["hist"] = list(hist)
        d["fault"] = rc
        return d

    def close(self):
        if self.h:
            lib().cubbs_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- primitive entries
def _res(n):
    return (Result * max(1, n))()


def _outs(keep, n, cap=48):
    """n output buffers of `cap` bytes: (array of ByteBuffer, list of ctypes buffers)."""
    arr = (ByteBuffer * max(1, n))()
    bufs = []
    for i in range(n):
        buf, bb = keep.out(cap)
        arr[i] = bb
        bufs.append(buf)
    keep.refs.append(arr)
    return arr, bufs


def _r(res, i):
    return (res[i].code, res[i].value, res[i].bytes_written)


def _msm_items(items, keep):
    """G1MsmItem array for [(points48, scalars32)] and the output buffers."""
    arr = (G1MsmItem * max(1, len(items)))()
    bufs = []
    for i, (pts, scs) in enumerate(items):
        arr[i].points = keep.slices(pts)
        arr[i].scalars = keep.slices(scs)
        arr[i].term_count = len(pts)
        buf, bb = keep.out(48)
        arr[i].point_out = bb
        bufs.append(buf)
    keep.refs.append(arr)
    return arr, bufs


def _common_items(groups, keep, packed, outputs=True):
    """CommonPointsItem array for [(points48, [row scalars])] and the row output buffers."""
    arr = (CommonPointsItem * max(1, len(groups)))()
    allbufs = []
    for g, (pts, rows) in enumerate(groups):
        arr[g].points = keep.slices(pts)
        arr[g].term_count = len(pts)
        if packed:
            flat = b"".join(b"".join(r) for r in rows)
            buf = ctypes.create_string_buffer(flat, max(1, len(flat)))
            keep.refs.append(buf)
            arr[g].scalar_rows_be = ctypes.cast(buf, ctypes.c_void_p)
            arr[g].scalar_row_stride = 32 * len(pts)
        else:
            arr[g].scalar_rows = keep.slices([s for r in rows for s in r])
        arr[g].output_count = len(rows)
        if outputs:
            outs, bufs = _outs(keep, len(rows))
            arr[g].point_outs = ctypes.cast(outs, ctypes.c_void_p)
            allbufs += bufs
    keep.refs.append(arr)
    return arr, allbufs


def g1_msm(items):
    """items: [(points48, scalars32)] -> (status, [(code, route, point or None)])."""
    keep = _Keep()
    arr, bufs = _msm_items(items, keep)
    res = _res(len(items))
    rc = lib().cubbs_bbs_gpu_g1_msm_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(items)), res)
    return rc, [(res[i].code, res[i].value, bufs[i].raw[:48] if res[i].code == 0 else None)
                for i in range(len(items))]


def g1_msm_common_points(groups, fused=False, packed=False):
    """groups: [(points48, [row scalars])] -> (status, [(code, route, point or None)] over all rows)."""
    keep = _Keep()
    arr, allbufs = _common_items(groups, keep, packed)
    res = _res(len(allbufs))
    fn = (lib().cubbs_bbs_gpu_g1_msm_fused_common_points_batch_v2 if fused
          else lib().cubbs_bbs_gpu_g1_msm_common_points_batch_v2)
    rc = fn(ABI_VERSION, arr, ctypes.c_size_t(len(groups)), res)
    return rc, [(res[i].code, res[i].value, allbufs[i].raw[:48] if res[i].code == 0 else None)
                for i in range(len(allbufs))]


def g1_msm_plus_variable(groups):
    """groups: [(points48, [(row scalars, var_point, var_scalar)])] -> (status, rows)."""
    keep = _Keep()
    arr = (PlusVariableItem * max(1, len(groups)))()
    allbufs = []
    for g, (pts, rows) in enumerate(groups):
        arr[g].points = keep.slices(pts)
        arr[g].term_count = len(pts)
        arr[g].scalar_rows = keep.slices([s for r, _, _ in rows for s in r])
        arr[g].variable_points = keep.slices([vp for _, vp, _ in rows])
        arr[g].variable_scalars = keep.slices([vs for _, _, vs in rows])
        arr[g].output_count = len(rows)
        outs, bufs = _outs(keep, len(rows))
        arr[g].point_outs = ctypes.cast(outs, ctypes.c_void_p)
        allbufs += bufs
    res = _res(len(allbufs))
    rc = lib().cubbs_bbs_gpu_g1_msm_common_points_plus_variable_batch_v2(
        ABI_VERSION, arr, ctypes.c_size_t(len(groups)), res)
    return rc, [(res[i].code, res[i].value, allbufs[i].raw[:48] if res[i].code == 0 else None)
                for i in range(len(allbufs))]


def proof_verify_msm_pair(groups):
    """groups: [(common_points48, [(common_scalars, [Bbar, Abar, D], [c, e^, r1^, r3^])])]
    -> (status, [(code, route, bytes, t1, t2)])."""
    keep = _Keep()
    arr = (PvMsmPairItem * max(1, len(groups)))()
    t1b, t2b = [], []
    for g, (pts, rows) in enumerate(groups):
        arr[g].common_points = keep.slices(pts)
        arr[g].common_term_count = len(pts)
        arr[g].common_scalar_rows = keep.slices([s for r, _, _ in rows for s in r])
        arr[g].proof_points = keep.slices([p for _, pp, _ in rows for p in pp])
        arr[g].proof_scalars = keep.slices([x for _, _, ps in rows for x in ps])
        arr[g].output_count = len(rows)
        o1, b1 = _outs(keep, len(rows))
        o2, b2 = _outs(keep, len(rows))
        arr[g].t1_outs = ctypes.cast(o1, ctypes.c_void_p)
        arr[g].t2_outs = ctypes.cast(o2, ctypes.c_void_p)
        t1b += b1
        t2b += b2
    res = _res(len(t1b))
    rc = lib().cubbs_bbs_gpu_proof_verify_msm_pair_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(groups)), res)
    return rc, [(res[i].code, res[i].value, res[i].bytes_written,
                 t1b[i].raw[:48] if res[i].code == 0 else None, t2b[i].raw[:48] if res[i].code == 0 else None)
                for i in range(len(t1b))]


def g1_dual_msm(items, lanes=None):
    """items: [(points, lhs_scalars, rhs_scalars)] -> (status, [(code, value, bytes, lhs, rhs)]);
    with lanes, each output is a list of lane sums."""
    keep = _Keep()
    n = len(items)
    arr = (DualMsmLanesItem * max(1, n))() if lanes else (DualMsmItem * max(1, n))()
    bufs = []
    for i, (pts, ls, rs) in enumerate(items):
        arr[i].points = keep.slices(pts)
        arr[i].lhs_scalars = keep.slices(ls)
        arr[i].rhs_scalars = keep.slices(rs)
        arr[i].term_count = len(pts)
        k = lanes[i] if lanes else 1
        lb, lbb = keep.out(48 * k)
        rb, rbb = keep.out(48 * k)
        if lanes:
            arr[i].lane_count = k
            arr[i].lhs_outs, arr[i].rhs_outs = lbb, rbb
        else:
            arr[i].lhs_out, arr[i].rhs_out = lbb, rbb
        bufs.append((lb, rb, k))
    res = _res(n)
    fn = (lib().cubbs_bbs_gpu_g1_dual_msm_same_points_lanes_batch_v2 if lanes
          else lib().cubbs_bbs_gpu_g1_dual_msm_same_points_batch_v2)
    rc = fn(ABI_VERSION, arr, ctypes.c_size_t(n), res)
    out = []
    for i, (lb, rb, k) in enumerate(bufs):
        if res[i].code:
            out.append((res[i].code, res[i].value, res[i].bytes_written, None, None))
            continue
        L = [lb.raw[48 * j:48 * j + 48] for j in range(k)]
        R = [rb.raw[48 * j:48 * j + 48] for j in range(k)]
        out.append((res[i].code, res[i].value, res[i].bytes_written,
                    L if lanes else L[0], R if lanes else R[0]))
    return rc, out


def g1_add(pairs):
    """pairs: [(a48, b48)] -> (status, [(code, point or None)])."""
    keep = _Keep()
    arr = (G1AddItem * max(1, len(pairs)))()
    bufs = []
    for i, (a, b) in enumerate(pairs):
        arr[i].lhs = keep.slice(a)
        arr[i].rhs = keep.slice(b)
        buf, bb = keep.out(48)
        arr[i].point_out = bb
        bufs.append(buf)
    res = _res(len(pairs))
    rc = lib().cubbs_bbs_gpu_g1_add_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(pairs)), res)
    return rc, [(res[i].code, bufs[i].raw[:48] if res[i].code == 0 else None) for i in range(len(pairs))]


def pairing_eq(items):
    """items: [(lhs_g1, lhs_g2, rhs_g1, rhs_g2)] -> (status, [(code, equal)])."""
    keep = _Keep()
    arr = (PairingEqItem * max(1, len(items)))()
    for i, (a1, a2, b1, b2) in enumerate(items):
        arr[i] = PairingEqItem(keep.slice(a1), keep.slice(a2), keep.slice(b1), keep.slice(b2))
    res = _res(len(items))
    rc = lib().cubbs_bbs_gpu_pairing_eq_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(items)), res)
    return rc, [(res[i].code, bool(res[i].value)) for i in range(len(items))]


def pairing_eq_aggregate(items):
    """items: [(lhs_g1, lhs_g2, rhs_g1, rhs_g2, coefficient32)] -> (status, code, equal)."""
    keep = _Keep()
    arr = (PairingEqScaledItem * max(1, len(items)))()
    for i, (a1, a2, b1, b2, c) in enumerate(items):
        arr[i] = PairingEqScaledItem(keep.slice(a1), keep.slice(a2), keep.slice(b1), keep.slice(b2),
                                     keep.slice(c))
    res = Result()
    rc = lib().cubbs_bbs_gpu_pairing_eq_aggregate_v2(ABI_VERSION, arr, ctypes.c_size_t(len(items)),
                                                     ctypes.byref(res))
    return rc, res.code, bool(res.value)


def prepare_b(items):
    """items: [(pk96, header64, [message scalars])] -> (status, [(code, B or None)])."""
    keep = _Keep()
    arr = (PrepareBItem * max(1, len(items)))()
    bufs = []
    for i, (pk, hdr, scs) in enumerate(items):
        arr[i].public_key = keep.slice(pk)
        arr[i].header = keep.slice(hdr)
        arr[i].message_scalars = keep.slices(scs)
        arr[i].message_count = len(scs)
        buf, bb = keep.out(48)
        arr[i].b_out = bb
        bufs.append(buf)
    res = _res(len(items))
    rc = lib().cubbs_bbs_gpu_prepare_b_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(items)), res)
    return rc, [(res[i].code, bufs[i].raw[:48] if res[i].code == 0 else None) for i in range(len(items))]


def sign_e_from_resident(items, message_dst, e_dst):
    """items: [(sk32, domain32, messages or None, scalar_offset, count)] ->
    (status, [(code, e, message_scalars or None)])."""
    keep = _Keep()
    arr = (SignEItem * max(1, len(items)))()
    bufs = []
    for i, (sk, dom, msgs, off, count) in enumerate(items):
        arr[i].secret_key = keep.slice(sk)
        arr[i].domain = keep.slice(dom)
        arr[i].messages = keep.slices(msgs) if msgs is not None else None
        arr[i].message_scalar_offset = off
        arr[i].message_count = count
        mb, mbb = keep.out(32 * count)
        eb, ebb = keep.out(32)
        arr[i].message_scalars_out = mbb
        arr[i].scalar_out = ebb
        bufs.append((eb, mb))
    res = _res(len(items))
    rc = lib().cubbs_bbs_gpu_sign_e_from_resident_scalars_batch_v2(
        ABI_VERSION, arr, ctypes.c_size_t(len(items)), keep.slice(message_dst), keep.slice(e_dst), res)
    return rc, [(res[i].code, bufs[i][0].raw[:32] if res[i].code == 0 else None,
                 bufs[i][1].raw[:32 * items[i][4]] if res[i].code == 0 and items[i][2] is not None else None)
                for i in range(len(items))]


def proof_gen_t2_prepare(items):
    """items: [(B48, A48, [7 scalars32])] -> (status, [(code, route, (d, abar, bbar, t1) or None)])."""
    keep = _Keep()
    arr = (T2PrepareItem * max(1, len(items)))()
    bufs = []
    for i, (b, a, scs) in enumerate(items):
        arr[i].compressed_b = keep.slice(b)
        arr[i].signature_a = keep.slice(a)
        for k in range(7):
            arr[i].scalars_be[k] = keep.slice(scs[k])
        outs = []
        for name in ("d_out", "abar_out", "bbar_out", "t1_out"):
            buf, bb = keep.out(48)
            setattr(arr[i], name, bb)
            outs.append(buf)
        bufs.append(outs)
    res = _res(len(items))
    rc = lib().cubbs_bbs_gpu_proof_gen_t2_prepare_batch_v2(ABI_VERSION, arr, ctypes.c_size_t(len(items)), res)
    return rc, [(res[i].code, res[i].value,
                 tuple(x.raw[:48] for x in bufs[i]) if res[i].code == 0 else None) for i in range(len(items))]
