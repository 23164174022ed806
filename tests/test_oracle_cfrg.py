"""The oracle pinned against the CFRG BBS vectors (bls12-381-sha-256) and
against closed forms.  Mirrors the reference protocol tests:
crates/cubbs-bbs-protocol/src/bbsplus/signature.rs:800-1015 (msg_signature),
proof.rs:3149-3560 (mocked_rng, proof_check), generators.rs / keys.rs /
utils/message.rs tests.  CPU only."""
import hashlib
import random

import pytest

from conftest import R_ORDER, load_json, proof_fixtures, signature_fixtures, trace_scalars


def test_sha256_fips_vectors(oracle):
    # FIPS 180-4 examples plus block-boundary lengths, against hashlib
    for m in [b"", b"abc", b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq", b"a" * 55,
              b"a" * 56, b"a" * 63, b"a" * 64, b"a" * 65, bytes(range(256)) * 5]:
        assert oracle.sha256(m) == hashlib.sha256(m).digest()


def _xmd_python(msg, dst, n):
    # RFC 9380 section 5.3.1 written out independently with hashlib
    ell = (n + 31) // 32
    dp = dst + bytes([len(dst)])
    b0 = hashlib.sha256(b"\x00" * 64 + msg + n.to_bytes(2, "big") + b"\x00" + dp).digest()
    out = [hashlib.sha256(b0 + b"\x01" + dp).digest()]
    for i in range(2, ell + 1):
        x = bytes(a ^ b for a, b in zip(b0, out[-1]))
        out.append(hashlib.sha256(x + bytes([i]) + dp).digest())
    return b"".join(out)[:n]


def test_expand_message_xmd_closed_form(oracle):
    rng = random.Random(7)
    for n in (32, 48, 64, 128, 480):
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(0, 300)))
        dst = b"QUUX-V01-CS02-with-expander-SHA256-128"
        assert oracle.expand_message_xmd(msg, dst, n) == _xmd_python(msg, dst, n)


def test_rfc9380_expand_message_xmd_vector(oracle):
    # RFC 9380 appendix K.1, msg = "" , len_in_bytes = 0x20
    dst = b"QUUX-V01-CS02-with-expander-SHA256-128"
    out = oracle.expand_message_xmd(b"", dst, 32)
    assert out.hex() == "68a985b87eb6b46952128911f2a4412bbc302a9d759667f87f7a21d803f07235"


def test_hash_to_scalar_fixture(oracle):
    d = load_json("h2s.json")
    assert oracle.hash_to_scalar(bytes.fromhex(d["message"]), bytes.fromhex(d["dst"])).hex() == d["scalar"]


def test_hash_to_scalar_is_okm_mod_r(oracle):
    # OS2IP(expand_message_xmd(msg, dst, 48)) mod r
    rng = random.Random(3)
    for _ in range(20):
        msg = bytes(rng.randrange(256) for _ in range(rng.randrange(0, 100)))
        okm = _xmd_python(msg, oracle.H2S_DST, 48)
        want = int.from_bytes(okm, "big") % R_ORDER
        assert int.from_bytes(oracle.hash_to_scalar(msg, oracle.H2S_DST), "big") == want


def test_map_message_to_scalar_fixture(oracle):
    d = load_json("MapMessageToScalarAsHash.json")
    assert bytes.fromhex(d["dst"]) == oracle.MAP_DST
    got = oracle.map_messages_to_scalars([bytes.fromhex(c["message"]) for c in d["cases"]])
    assert [g.hex() for g in got] == [c["scalar"] for c in d["cases"]]


def test_mocked_random_scalars_fixture(oracle):
    d = load_json("mockedRng.json")
    got = oracle.seeded_random_scalars(bytes.fromhex(d["seed"]), bytes.fromhex(d["dst"]), d["count"])
    assert [g.hex() for g in got] == d["mockedScalars"]


def test_generators_fixture(oracle):
    d = load_json("generators.json")
    gens = oracle.generators(1 + len(d["MsgGenerators"]))
    assert oracle.p1().hex() == d["P1"]
    assert gens[0].hex() == d["Q1"]
    assert [g.hex() for g in gens[1:]] == d["MsgGenerators"]
    # every generator is a valid, prime-order-subgroup point
    assert all(oracle.g1_valid(g) for g in gens)


def test_keypair_fixture(oracle):
    d = load_json("keypair.json")
    sk, pk = oracle.keygen(bytes.fromhex(d["keyMaterial"]), bytes.fromhex(d["keyInfo"]))
    assert sk.hex() == d["keyPair"]["secretKey"]
    assert pk.hex() == d["keyPair"]["publicKey"]


@pytest.mark.parametrize("fx", signature_fixtures(), ids=lambda d: d["caseName"][:40])
def test_signature_fixture(oracle, fx):
    sk = bytes.fromhex(fx["signerKeyPair"]["secretKey"])
    pk = bytes.fromhex(fx["signerKeyPair"]["publicKey"])
    hdr = bytes.fromhex(fx["header"])
    msgs = [bytes.fromhex(m) for m in fx["messages"]]
    sig = bytes.fromhex(fx["signature"])
    assert oracle.verify(pk, sig, hdr, msgs) == fx["result"]["valid"]
    if fx["result"]["valid"]:
        assert oracle.sign(sk, pk, hdr, msgs) == sig
        assert oracle.compute_b(pk, hdr, msgs).hex() == fx["trace"]["B"]
        assert oracle.domain(pk, len(msgs), hdr).hex() == fx["trace"]["domain"]


@pytest.mark.parametrize("fx", proof_fixtures(), ids=lambda d: d["caseName"][:60])
def test_proof_fixture(oracle, fx):
    pk = bytes.fromhex(fx["signerPublicKey"])
    sig = bytes.fromhex(fx["signature"])
    hdr = bytes.fromhex(fx["header"])
    ph = bytes.fromhex(fx["presentationHeader"])
    msgs = [bytes.fromhex(m) for m in fx["messages"]]
    idx = fx["disclosedIndexes"]
    proof = bytes.fromhex(fx["proof"])
    dmsgs = [msgs[i] for i in idx if i < len(msgs)]
    assert oracle.proof_verify(pk, proof, hdr, ph, dmsgs, idx, len(msgs)) == fx["result"]["valid"]
    if fx["result"]["valid"]:
        got, tr = oracle.proof_gen(pk, sig, hdr, ph, msgs, idx, trace_scalars(fx["trace"]), with_trace=True)
        assert got == proof
        for k in ("A_bar", "B_bar", "D", "T1", "T2", "domain", "challenge"):
            assert tr[k].hex() == fx["trace"][k], k


def test_proof_fixture_random_scalars_are_mocked_rng(oracle):
    # the traces' blinding scalars are seeded_random_scalars(SEED, MOCK DST)
    # (proof.rs:1614-1615): pins the scalar order [r1, r2, e~, r1~, r3~, m~..]
    d = load_json("mockedRng.json")
    fx = load_json("proof/proof003.json")
    rs = trace_scalars(fx["trace"])
    got = oracle.seeded_random_scalars(bytes.fromhex(d["seed"]), bytes.fromhex(d["dst"]), len(rs))
    assert got == rs


def test_pairing_bilinear_and_nondegenerate(oracle):
    rng = random.Random(11)
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    for _ in range(2):
        a, b = rng.randrange(1, R_ORDER), rng.randrange(1, R_ORDER)
        ab = a * b % R_ORDER
        lhs = oracle.pairing_gt(oracle.g1_mul(g1, a.to_bytes(32, "big")), oracle.g2_mul(g2, b.to_bytes(32, "big")))
        rhs = oracle.pairing_gt(oracle.g1_mul(g1, ab.to_bytes(32, "big")), g2)
        assert lhs == rhs
        one = bytes(47) + b"\x01" + bytes(576 - 48)
        assert lhs != one
    # e(P, Q) e(-P, Q) == 1 via the equality form e(P,Q) == e(P,Q)
    assert oracle.pairing_eq(g1, g2, g1, g2) == 1
    assert oracle.pairing_eq(g1, g2, oracle.g1_mul(g1, (2).to_bytes(32, "big")), g2) == 0


def test_final_exponentiation_matches_plain_power(oracle):
    # x-chain hard part == cube of the plain (p^12-1)/r exponentiation
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    assert oracle.check_final_exp(g1, g2)
    p = oracle.g1_mul(g1, (123456789).to_bytes(32, "big"))
    assert oracle.check_final_exp(p, g2)


def test_curve_parameters_from_x():
    x = -0xD201000000010000
    p = ((x - 1) ** 2 * (x ** 4 - x ** 2 + 1)) // 3 + x
    assert p == 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
    assert x ** 4 - x ** 2 + 1 == R_ORDER


def test_chacha20_rfc8439_block(oracle):
    # RFC 8439 section 2.3.2 test vector
    key = bytes(range(32))
    nonce = bytes.fromhex("000000090000004a00000000")
    out = oracle.chacha20_block(key, 1, nonce)
    assert out.hex() == (
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")


def test_chacha_scalars_convention(oracle):
    seed = hashlib.sha256(b"cubbs-hip/config4").digest()
    got = oracle.chacha_random_scalars(seed, 3)
    for k, s in enumerate(got):
        blk = oracle.chacha20_block(seed, k, bytes(12))
        assert int.from_bytes(s, "big") == int.from_bytes(blk, "little") % R_ORDER


def test_point_codec_rejections(oracle):
    g1 = oracle.g1_generator()
    assert oracle.g1_valid(g1)
    bad = bytearray(g1)
    bad[0] &= 0x7F  # compression flag cleared
    assert not oracle.g1_valid(bytes(bad))
    assert oracle.g1_valid(b"\xc0" + bytes(47))         # identity encoding
    assert not oracle.g1_valid(b"\xe0" + bytes(47))     # identity with sign bit
    assert not oracle.g1_valid(b"\xc0" + bytes(46) + b"\x01")
    # x >= p
    p = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
    xp = bytearray(p.to_bytes(48, "big"))
    xp[0] |= 0x80
    assert not oracle.g1_valid(bytes(xp))
    # a point on the curve but outside G1 (cofactor): x with x^3 + 4 a square
    for x in range(1, 200):
        rhs = (x ** 3 + 4) % p
        y = pow(rhs, (p + 1) // 4, p)
        if y * y % p == rhs:
            enc = bytearray(x.to_bytes(48, "big"))
            enc[0] |= 0x80
            assert not oracle.g1_valid(bytes(enc))
            break


def test_max_messages_fixture_inputs(oracle):
    """tests/golden/max_messages/l65536.json: the key comes from its IKM and
    the header follows the reference's 65,536-message test
    (crates/cubbs-bbs/src/tests.rs:919-921); the signature's e is canonical."""
    import json
    import os
    from conftest import ROOT
    v = json.load(open(os.path.join(ROOT, "tests", "golden", "max_messages", "l65536.json")))
    sk, pk = oracle.keygen(bytes.fromhex(v["ikm"]))
    assert (sk.hex(), pk.hex()) == (v["sk"], v["pk"])
    hdr = bytes.fromhex(v["header"])
    assert len(hdr) == 64 and hdr[:8] == v["L"].to_bytes(8, "big") and hdr[8:33] == b"omniid-route4-correctness"
    assert hdr[33:] == bytes(31)
    sig = bytes.fromhex(v["signature"])
    assert len(sig) == 80 and oracle.g1_valid(sig[:48])
    assert int.from_bytes(sig[48:], "big") < R_ORDER
    assert v["proof_len"] == 272 + 32 * (v["L"] // 2)
