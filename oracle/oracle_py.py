"""ctypes binding of the CPU oracle (test infrastructure only).

Only the tests, the smoke check and the benchmark's CPU-baseline leg import
this module.  The HIP product never does.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bbs.so")
_lib = None

API_ID = b"BBS_BLS12381G1_XMD:SHA-256_SSWU_RO_H2G_HM2S_"
MAP_DST = API_ID + b"MAP_MSG_TO_SCALAR_AS_HASH_"
H2S_DST = API_ID + b"H2S_"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_init()
        _lib.orc_proof_gen.restype = ctypes.c_long
    return _lib


def _buf(b):
    return ctypes.create_string_buffer(bytes(b), len(b) or 1)


def _pack(msgs):
    data = b"".join(msgs)
    lens = (ctypes.c_size_t * max(1, len(msgs)))(*[len(m) for m in msgs])
    return _buf(data), lens


def _sizes(vals):
    return (ctypes.c_size_t * max(1, len(vals)))(*vals)


def sha256(msg):
    out = ctypes.create_string_buffer(32)
    lib().orc_sha256(_buf(msg), ctypes.c_size_t(len(msg)), out)
    return out.raw


def expand_message_xmd(msg, dst, n):
    out = ctypes.create_string_buffer(n)
    rc = lib().orc_expand_message_xmd(_buf(msg), ctypes.c_size_t(len(msg)), _buf(dst),
                                      ctypes.c_size_t(len(dst)), out, ctypes.c_size_t(n))
    assert rc == 0
    return out.raw


def hash_to_scalar(msg, dst):
    out = ctypes.create_string_buffer(32)
    rc = lib().orc_hash_to_scalar(_buf(msg), ctypes.c_size_t(len(msg)), _buf(dst),
                                  ctypes.c_size_t(len(dst)), out)
    assert rc == 0
    return out.raw


def map_messages_to_scalars(msgs):
    out = ctypes.create_string_buffer(32 * max(1, len(msgs)))
    data, lens = _pack(msgs)
    lib().orc_map_messages_to_scalars(data, lens, ctypes.c_size_t(len(msgs)), out)
    return [out.raw[32 * i:32 * i + 32] for i in range(len(msgs))]


def seeded_random_scalars(seed, dst, count):
    out = ctypes.create_string_buffer(32 * max(1, count))
    rc = lib().orc_seeded_random_scalars(_buf(seed), ctypes.c_size_t(len(seed)), _buf(dst),
                                         ctypes.c_size_t(len(dst)), ctypes.c_size_t(count), out)
    assert rc == 0
    return [out.raw[32 * i:32 * i + 32] for i in range(count)]


def chacha_random_scalars(seed, count):
    out = ctypes.create_string_buffer(32 * max(1, count))
    lib().orc_chacha_random_scalars(_buf(seed), ctypes.c_size_t(count), out)
    return [out.raw[32 * i:32 * i + 32] for i in range(count)]


def chacha20_block(key, counter, nonce):
    out = ctypes.create_string_buffer(64)
    lib().orc_chacha20_block(_buf(key), ctypes.c_uint32(counter), _buf(nonce), out)
    return out.raw


def g1_valid(p):
    return lib().orc_g1_decompress_check(_buf(p)) == 1


def g2_valid(q):
    return lib().orc_g2_decompress_check(_buf(q)) == 1


def g1_add(a, b):
    out = ctypes.create_string_buffer(48)
    rc = lib().orc_g1_add(_buf(a), _buf(b), out)
    return out.raw if rc == 0 else None


def g1_mul(p, k):
    out = ctypes.create_string_buffer(48)
    rc = lib().orc_g1_mul(_buf(p), _buf(k), out)
    return out.raw if rc == 0 else None


def g1_msm(points, scalars):
    out = ctypes.create_string_buffer(48)
    rc = lib().orc_g1_msm(_buf(b"".join(points)), _buf(b"".join(scalars)),
                          ctypes.c_size_t(len(points)), out)
    return out.raw if rc == 0 else None


def g1_generator():
    out = ctypes.create_string_buffer(48)
    lib().orc_g1_generator(out)
    return out.raw


def g2_generator():
    out = ctypes.create_string_buffer(96)
    lib().orc_g2_generator(out)
    return out.raw


def g2_mul(q, k):
    out = ctypes.create_string_buffer(96)
    rc = lib().orc_g2_mul(_buf(q), _buf(k), out)
    return out.raw if rc == 0 else None


def hash_to_curve_g1(msg, dst):
    out = ctypes.create_string_buffer(48)
    lib().orc_hash_to_curve_g1(_buf(msg), ctypes.c_size_t(len(msg)), _buf(dst),
                               ctypes.c_size_t(len(dst)), out)
    return out.raw


def pairing_eq(a1, a2, b1, b2):
    return lib().orc_pairing_eq(_buf(a1), _buf(a2), _buf(b1), _buf(b2))


def pairing_gt(p, q, naive=False):
    out = ctypes.create_string_buffer(576)
    rc = lib().orc_pairing_gt(_buf(p), _buf(q), int(naive), out)
    assert rc == 0
    return out.raw


def fr_mul(a, b):
    out = ctypes.create_string_buffer(32)
    assert lib().orc_fr_mul(_buf(a), _buf(b), out) == 0
    return out.raw


def fr_add(a, b):
    out = ctypes.create_string_buffer(32)
    assert lib().orc_fr_add(_buf(a), _buf(b), out) == 0
    return out.raw


def fr_inv(a):
    out = ctypes.create_string_buffer(32)
    assert lib().orc_fr_inv(_buf(a), out) == 0
    return out.raw


def keygen(ikm, info=b""):
    sk = ctypes.create_string_buffer(32)
    pk = ctypes.create_string_buffer(96)
    rc = lib().orc_keygen(_buf(ikm), ctypes.c_size_t(len(ikm)), _buf(info),
                          ctypes.c_size_t(len(info)), sk, pk)
    assert rc == 0
    return sk.raw, pk.raw


def sk_to_pk(sk):
    pk = ctypes.create_string_buffer(96)
    assert lib().orc_sk_to_pk(_buf(sk), pk) == 0
    return pk.raw


def generators(count):
    out = ctypes.create_string_buffer(48 * count)
    lib().orc_generators(ctypes.c_size_t(count), out)
    return [out.raw[48 * i:48 * i + 48] for i in range(count)]


def p1():
    out = ctypes.create_string_buffer(48)
    lib().orc_p1(out)
    return out.raw


def domain(pk, L, header):
    out = ctypes.create_string_buffer(32)
    lib().orc_domain(_buf(pk), ctypes.c_size_t(L), _buf(header), ctypes.c_size_t(len(header)), out)
    return out.raw


def compute_b(pk, header, msgs):
    out = ctypes.create_string_buffer(48)
    data, lens = _pack(msgs)
    lib().orc_compute_b(_buf(pk), _buf(header), ctypes.c_size_t(len(header)), data, lens,
                        ctypes.c_size_t(len(msgs)), out)
    return out.raw


def sign(sk, pk, header, msgs):
    out = ctypes.create_string_buffer(80)
    data, lens = _pack(msgs)
    rc = lib().orc_sign(_buf(sk), _buf(pk), _buf(header), ctypes.c_size_t(len(header)), data, lens,
                        ctypes.c_size_t(len(msgs)), out)
    return out.raw if rc == 0 else None


def verify(pk, sig, header, msgs):
    """True / False, or None on an input error (the reference's Err)."""
    data, lens = _pack(msgs)
    rc = lib().orc_verify(_buf(pk), ctypes.c_size_t(len(pk)), _buf(sig), ctypes.c_size_t(len(sig)),
                          _buf(header), ctypes.c_size_t(len(header)), data, lens,
                          ctypes.c_size_t(len(msgs)))
    return None if rc < 0 else bool(rc)


def proof_gen(pk, sig, header, ph, msgs, disclosed, random_scalars, with_trace=False):
    L = len(msgs)
    U = L - len(disclosed)
    cap = 272 + 32 * max(U, 0)
    out = ctypes.create_string_buffer(cap)
    trace = ctypes.create_string_buffer(304)
    data, lens = _pack(msgs)
    rc = lib().orc_proof_gen(_buf(pk), _buf(sig), ctypes.c_size_t(len(sig)), _buf(header),
                             ctypes.c_size_t(len(header)), _buf(ph), ctypes.c_size_t(len(ph)), data,
                             lens, ctypes.c_size_t(L), _sizes(disclosed),
                             ctypes.c_size_t(len(disclosed)), _buf(b"".join(random_scalars)), out,
                             ctypes.c_size_t(cap), trace)
    if rc < 0:
        return None
    if with_trace:
        t = trace.raw
        return out.raw[:rc], {"A_bar": t[0:48], "B_bar": t[48:96], "D": t[96:144],
                              "T1": t[144:192], "T2": t[192:240], "domain": t[240:272],
                              "challenge": t[272:304]}
    return out.raw[:rc]


def proof_verify(pk, proof, header, ph, disclosed_msgs, disclosed, total):
    data, lens = _pack(disclosed_msgs)
    rc = lib().orc_proof_verify(_buf(pk), ctypes.c_size_t(len(pk)), _buf(proof),
                                ctypes.c_size_t(len(proof)), _buf(header),
                                ctypes.c_size_t(len(header)), _buf(ph), ctypes.c_size_t(len(ph)),
                                data, lens, ctypes.c_size_t(len(disclosed_msgs)), _sizes(disclosed),
                                ctypes.c_size_t(len(disclosed)), ctypes.c_size_t(total))
    return None if rc < 0 else bool(rc)


def check_final_exp(p, q):
    return lib().orc_check_final_exp(_buf(p), _buf(q)) == 1
