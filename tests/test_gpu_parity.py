"""GPU parity tests: every call goes through libcubbs_bbs_hip.so (the C ABI v2
and its extensions) and is compared with the CPU oracle on the same seeded
inputs, or with the CFRG golden vectors.  Bit-exact for bytes, equal for
booleans and per-item result codes.

The matrix follows the reference's own tests: protocol fixtures
(crates/cubbs-bbs-protocol/src/bbsplus/{signature,proof}.rs tests), the native
smoke tests (native/cuda/tests/{sha256,hash_to_scalar,prepare_b}_smoke.cpp) and
the engine's GPU-vs-CPU matrix (crates/cubbs-bbs/src/tests.rs:466-1313: mixed
message counts, shared/unique issuers, tampering, empty and large inputs).
"""
import hashlib
import os
import random

import pytest

from conftest import R_ORDER, load_json, proof_fixtures, signature_fixtures, trace_scalars

pytestmark = pytest.mark.gpu

P_MOD = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB


def rnd_bytes(rng, n):
    return bytes(rng.randrange(256) for _ in range(n))


# ------------------------------------------------------------------ hashing
def test_backend_info(gpu):
    rc, kind, hw = gpu.backend_info()
    assert (rc, kind, hw) == (0, 1, 1)


def test_sha256_smoke_vectors(gpu):
    # native/cuda/tests/sha256_smoke.cpp style: FIPS messages and block edges
    msgs = [b"", b"abc", b"abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq", b"a" * 55,
            b"a" * 56, b"a" * 64, b"a" * 119, bytes(range(256)) * 3]
    rc, out = gpu.sha256_batch(msgs)
    assert rc == 0
    assert out == [hashlib.sha256(m).digest() for m in msgs]


def test_expand_xmd_and_h2s_match_oracle(gpu, oracle):
    rng = random.Random(5)
    msgs = [rnd_bytes(rng, rng.randrange(0, 300)) for _ in range(64)]
    for dst in (oracle.MAP_DST, oracle.H2S_DST, b"", b"x" * 255):
        rc, x = gpu.expand_xmd48_batch(msgs, dst)
        assert rc == 0
        assert x == [oracle.expand_message_xmd(m, dst, 48) for m in msgs]
        rc, s = gpu.hash_to_scalar_batch(msgs, dst)
        assert rc == 0
        assert s == [oracle.hash_to_scalar(m, dst) for m in msgs]


def test_hash_to_scalar_fixtures(gpu):
    d = load_json("h2s.json")
    rc, s = gpu.hash_to_scalar_batch([bytes.fromhex(d["message"])], bytes.fromhex(d["dst"]))
    assert rc == 0 and s[0].hex() == d["scalar"]
    m = load_json("MapMessageToScalarAsHash.json")
    rc, sc = gpu.messages_to_scalars([bytes.fromhex(c["message"]) for c in m["cases"]])
    assert rc == 0
    assert [x.hex() for x in sc] == [c["scalar"] for c in m["cases"]]


# ------------------------------------------------------------------ curve
def test_generators_fixture_and_oracle(gpu, oracle):
    d = load_json("generators.json")
    rc, gens = gpu.generators(160)
    assert rc == 0
    assert gens[0].hex() == d["P1"] and gens[1].hex() == d["Q1"]
    assert [g.hex() for g in gens[2:12]] == d["MsgGenerators"]
    # beyond the fixture: same chain and hash_to_curve as the oracle
    assert gens[1:] == oracle.generators(159)


def test_g1_glv_mul_matches_oracle(gpu, oracle):
    rng = random.Random(21)
    g = oracle.g1_generator()
    lam = (0xD201000000010000 ** 2 - 1) % R_ORDER
    edge = [0, 1, 2, R_ORDER - 1, R_ORDER - 2, lam, lam - 1, lam + 1, 1 << 128, (1 << 254) + 7,
            lam * 2 % R_ORDER, (lam * lam) % R_ORDER]
    pts, scs = [], []
    for k in edge + [rng.randrange(R_ORDER) for _ in range(52)]:
        pts.append(oracle.g1_mul(g, rng.randrange(1, R_ORDER).to_bytes(32, "big")))
        scs.append(k.to_bytes(32, "big"))
    pts.append(b"\xc0" + bytes(47))  # identity base
    scs.append((12345).to_bytes(32, "big"))
    rc, out = gpu.g1_mul(pts, scs)
    assert rc == 0
    assert out == [oracle.g1_mul(p, s) for p, s in zip(pts, scs)]


def test_g1_decoding_rejections(gpu, oracle):
    g = oracle.g1_generator()
    one = (1).to_bytes(32, "big")
    bad = []
    b = bytearray(g)
    b[0] &= 0x7F
    bad.append(bytes(b))                     # not compressed
    bad.append(b"\xe0" + bytes(47))          # identity with sign flag
    bad.append(b"\xc0" + bytes(46) + b"\x01")
    xp = bytearray(P_MOD.to_bytes(48, "big"))
    xp[0] |= 0x80
    bad.append(bytes(xp))                    # x = p
    # points on the curve outside G1, and x with no curve point
    found_off, found_none = 0, 0
    for x in range(1, 400):
        rhs = (x ** 3 + 4) % P_MOD
        y = pow(rhs, (P_MOD + 1) // 4, P_MOD)
        enc = bytearray(x.to_bytes(48, "big"))
        enc[0] |= 0x80
        if y * y % P_MOD == rhs and found_off < 3:
            bad.append(bytes(enc))
            found_off += 1
        elif y * y % P_MOD != rhs and found_none < 3:
            bad.append(bytes(enc))
            found_none += 1
    for p in bad:
        assert not oracle.g1_valid(p)
    rc, out = gpu.g1_mul(bad, [one] * len(bad))
    assert rc == 0
    assert out == [None] * len(bad)
    # non-canonical scalar (= r) is rejected too (prepare_b_smoke.cpp:531-563)
    rc, out = gpu.g1_mul([g], [R_ORDER.to_bytes(32, "big")])
    assert out == [None]


def test_g2_codec_and_subgroup(gpu, oracle):
    rng = random.Random(4)
    g2 = oracle.g2_generator()
    qs = [g2] + [oracle.g2_mul(g2, rng.randrange(1, R_ORDER).to_bytes(32, "big")) for _ in range(5)]
    qs.append(b"\xc0" + bytes(95))
    rc, out = gpu.g2_roundtrip(qs)
    assert rc == 0 and out == qs
    # x = 5: either no point on the twist or a point outside G2
    enc = b"\x80" + bytes(94) + b"\x05"
    assert not oracle.g2_valid(enc)
    rc, out = gpu.g2_roundtrip([enc])
    assert out == [None]


def test_pairing_gt_bit_exact(gpu, oracle):
    rng = random.Random(8)
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    ps = [oracle.g1_mul(g1, rng.randrange(1, R_ORDER).to_bytes(32, "big")) for _ in range(3)]
    qs = [oracle.g2_mul(g2, rng.randrange(1, R_ORDER).to_bytes(32, "big")) for _ in range(3)]
    rc, out = gpu.pairing_gt(ps, qs)
    assert rc == 0
    assert out == [oracle.pairing_gt(p, q) for p, q in zip(ps, qs)]


# ------------------------------------------------------------------ keys and CFRG fixtures
def test_keygen_fixture(gpu):
    d = load_json("keypair.json")
    rc, kp = gpu.keygen([bytes.fromhex(d["keyMaterial"])], bytes.fromhex(d["keyInfo"]))
    assert rc == 0
    assert kp[0][0].hex() == d["keyPair"]["secretKey"]
    assert kp[0][1].hex() == d["keyPair"]["publicKey"]


def test_sign_verify_cfrg_fixtures(gpu):
    fx = signature_fixtures()
    reqs = [(bytes.fromhex(d["signerKeyPair"]["publicKey"]), bytes.fromhex(d["signature"]),
             bytes.fromhex(d["header"]), [bytes.fromhex(m) for m in d["messages"]]) for d in fx]
    rc, res = gpu.verify_batch(reqs)
    assert rc == 0
    assert res == [(0, d["result"]["valid"]) for d in fx]
    valid = [d for d in fx if d["result"]["valid"]]
    rc, sigs = gpu.sign_batch([(bytes.fromhex(d["signerKeyPair"]["secretKey"]),
                                bytes.fromhex(d["signerKeyPair"]["publicKey"]),
                                bytes.fromhex(d["header"]), [bytes.fromhex(m) for m in d["messages"]])
                               for d in valid])
    assert rc == 0
    assert sigs == [(0, bytes.fromhex(d["signature"])) for d in valid]


def _pv_req(d):
    msgs = [bytes.fromhex(m) for m in d["messages"]]
    idx = d["disclosedIndexes"]
    return (bytes.fromhex(d["signerPublicKey"]), bytes.fromhex(d["proof"]), bytes.fromhex(d["header"]),
            bytes.fromhex(d["presentationHeader"]), [msgs[i] for i in idx if i < len(msgs)], idx, len(msgs))


def test_proof_cfrg_fixtures(gpu, oracle):
    fx = proof_fixtures()
    reqs = [_pv_req(d) for d in fx]
    rc, res = gpu.proof_verify_batch(reqs)
    assert rc == 0
    assert res == [(0, d["result"]["valid"]) for d in fx]
    # ProofGen with the traces' blinding scalars is byte-exact
    valid = [d for d in fx if d["result"]["valid"]]
    greqs = [(bytes.fromhex(d["signerPublicKey"]), bytes.fromhex(d["signature"]), bytes.fromhex(d["header"]),
              bytes.fromhex(d["presentationHeader"]), [bytes.fromhex(m) for m in d["messages"]],
              d["disclosedIndexes"]) for d in valid]
    rc, proofs = gpu.proof_gen_batch(greqs, scalars=[trace_scalars(d["trace"]) for d in valid])
    assert rc == 0
    assert proofs == [(0, bytes.fromhex(d["proof"])) for d in valid]


# ------------------------------------------------------------------ seeded workloads vs oracle
def _issuers(gpu, rng, n):
    rc, kps = gpu.keygen([rnd_bytes(rng, 32) for _ in range(n)])
    assert rc == 0
    return kps


def test_config1_sign_then_verify(gpu, oracle):
    # one keypair from a seeded 32-byte IKM; 64 Sign then Verify calls, 10 messages of
    # 32 random bytes and a 16-byte header each; signatures byte-compared
    rng = random.Random(hashlib.sha256(b"cubbs-hip/config1").digest())
    ikm = rnd_bytes(rng, 32)
    sk, pk = oracle.keygen(ikm)
    rc, kp = gpu.keygen([ikm])
    assert kp[0] == (sk, pk)
    reqs = [(sk, pk, rnd_bytes(rng, 16), [rnd_bytes(rng, 32) for _ in range(10)]) for _ in range(64)]
    rc, sigs = gpu.sign_batch(reqs)
    assert rc == 0
    want = [oracle.sign(*r) for r in reqs]
    assert [s for _, s in sigs] == want
    rc, res = gpu.verify_batch([(pk, s, h, m) for (sk_, pk_, h, m), s in zip(reqs, want)])
    assert res == [(0, True)] * 64


def test_uneven_verify_mixed_issuers_and_tampering(gpu, oracle):
    rng = random.Random(99)
    kps = _issuers(gpu, rng, 5)
    reqs, expect = [], []
    for i in range(40):
        sk, pk = kps[rng.randrange(5)]
        L = rng.choice([0, 1, 2, 3, 7, 16, 33, 64, 128])
        msgs = [rnd_bytes(rng, rng.randrange(0, 257)) for _ in range(L)]
        hdr = rnd_bytes(rng, rng.choice([0, 16, 64]))
        sig = oracle.sign(sk, pk, hdr, msgs)
        kind = i % 5
        if kind == 1 and L:
            msgs = list(msgs)
            msgs[rng.randrange(L)] = rnd_bytes(rng, 5)
        elif kind == 2:
            hdr = hdr + b"!"
        elif kind == 3:
            pk = kps[(kps.index((sk, pk)) + 1) % 5][1]
        reqs.append((pk, sig, hdr, msgs))
        expect.append(oracle.verify(pk, sig, hdr, msgs))
    rc, res = gpu.verify_batch(reqs)
    assert rc == 0
    assert res == [(0, e) for e in expect]
    # each result is the same alone and in the mix
    for k in (0, 7, 13, 39):
        rc, alone = gpu.verify_batch([reqs[k]])
        assert alone[0] == res[k]


def test_verify_error_codes(gpu, oracle):
    rng = random.Random(12)
    (sk, pk), = _issuers(gpu, rng, 1)
    msgs = [b"a", b"b"]
    sig = oracle.sign(sk, pk, b"", msgs)
    bad_e = sig[:48] + R_ORDER.to_bytes(32, "big")             # non-canonical e
    bad_a = b"\x80" + bytes(46) + b"\x05" + sig[48:]            # A off the curve / subgroup
    bad_pk = b"\x80" + bytes(94) + b"\x05"
    id_pk = b"\xc0" + bytes(95)
    reqs = [(pk, sig, b"", msgs), (pk, bad_e, b"", msgs), (pk, bad_a, b"", msgs), (bad_pk, sig, b"", msgs),
            (pk[:95], sig, b"", msgs), (pk, sig[:79], b"", msgs), (id_pk, sig, b"", msgs),
            (pk, b"\xc0" + bytes(47) + sig[48:], b"", msgs)]
    rc, res = gpu.verify_batch(reqs)
    assert rc == 0
    codes = [c for c, _ in res]
    assert res[0] == (0, True)
    assert codes[1:6] == [1, 1, 1, 1, 1]
    # identity PK and identity A are valid encodings: the check just fails
    assert res[6] == (0, oracle.verify(*reqs[6]))
    assert res[7] == (0, oracle.verify(*reqs[7]))


def test_sign_error_codes(gpu, oracle):
    rng = random.Random(13)
    (sk, pk), = _issuers(gpu, rng, 1)
    reqs = [(sk, pk, b"h", [b"m"]), (R_ORDER.to_bytes(32, "big"), pk, b"h", [b"m"]),
            (sk, b"\x80" + bytes(94) + b"\x05", b"h", [b"m"]), (sk[:31], pk, b"h", [b"m"]),
            (sk, pk, b"", [])]
    rc, res = gpu.sign_batch(reqs)
    assert rc == 0
    assert res[0] == (0, oracle.sign(*reqs[0]))
    assert [c for c, _ in res[1:4]] == [1, 1, 1]
    assert res[4] == (0, oracle.sign(*reqs[4]))


def test_proof_gen_seeded_matches_oracle_and_verifies(gpu, oracle):
    rng = random.Random(31)
    kps = _issuers(gpu, rng, 3)
    greqs, seeds, pvs = [], [], []
    for i in range(12):
        sk, pk = kps[i % 3]
        L = rng.choice([1, 2, 5, 12, 32])
        msgs = [rnd_bytes(rng, rng.randrange(0, 100)) for _ in range(L)]
        hdr = rnd_bytes(rng, rng.choice([0, 64]))
        sig = oracle.sign(sk, pk, hdr, msgs)
        R = rng.randrange(0, L + 1)
        idx = sorted(rng.sample(range(L), R))
        ph = rnd_bytes(rng, rng.choice([0, 64]))
        greqs.append((pk, sig, hdr, ph, msgs, idx))
        seeds.append(rnd_bytes(rng, 32))
    rc, proofs = gpu.proof_gen_batch(greqs, seeds=seeds)
    assert rc == 0
    for (pk, sig, hdr, ph, msgs, idx), seed, (code, proof) in zip(greqs, seeds, proofs):
        assert code == 0
        U = len(msgs) - len(idx)
        want = oracle.proof_gen(pk, sig, hdr, ph, msgs, idx, oracle.chacha_random_scalars(seed, 5 + U))
        assert proof == want
        pvs.append((pk, proof, hdr, ph, [msgs[i] for i in idx], idx, len(msgs)))
    # verify all, plus tampered variants
    tampered = []
    for k, (pk, proof, hdr, ph, dm, idx, total) in enumerate(pvs[:6]):
        if k % 3 == 0:
            tampered.append((pk, proof, hdr, ph + b"x", dm, idx, total))
        elif k % 3 == 1:
            tampered.append((pk, proof, hdr + b"x", ph, dm, idx, total))
        else:
            p2 = bytearray(proof)
            p2[150] ^= 1
            tampered.append((pk, bytes(p2), hdr, ph, dm, idx, total))
    allreqs = pvs + tampered
    rc, res = gpu.proof_verify_batch(allreqs)
    assert rc == 0
    assert res == [(0, True)] * len(pvs) + [
        (0 if oracle.proof_verify(*t) is not None else 1, bool(oracle.proof_verify(*t))) for t in tampered]


def test_proof_verify_index_semantics(gpu, oracle):
    # protocol.rs:774-779: unsorted / duplicate / out-of-range indexes -> false;
    # proof.rs:3027-3033: an index beyond U + R - 1 -> false
    d = load_json("proof/proof003.json")
    pk, proof, hdr, ph, dm, idx, total = _pv_req(d)
    reqs = [(pk, proof, hdr, ph, dm, idx, total),
            (pk, proof, hdr, ph, dm[::-1], idx[::-1], total),
            (pk, proof, hdr, ph, dm, idx, idx[-1]),
            (pk, proof, hdr, ph, dm[:-1], idx, total),
            (pk, proof[:271], hdr, ph, dm, idx, total),
            (pk, proof + b"\x01\x02", hdr, ph, dm, idx, total)]
    rc, res = gpu.proof_verify_batch(reqs)
    assert rc == 0
    assert res[0] == (0, True)
    assert res[1] == (0, False) and res[2] == (0, False) and res[3] == (0, False)
    assert res[4][0] == 1
    # trailing bytes that do not fill a 32-byte chunk are ignored (proof.rs:206-213)
    assert res[5] == (0, True)
    assert oracle.proof_verify(*reqs[5]) is True


def test_large_message_count_grows_generators(gpu, oracle):
    rng = random.Random(77)
    (sk, pk), = _issuers(gpu, rng, 1)
    msgs = [rnd_bytes(rng, 8) for _ in range(300)]
    sig = oracle.sign(sk, pk, b"", msgs)
    rc, sigs = gpu.sign_batch([(sk, pk, b"", msgs)])
    assert sigs[0] == (0, sig)
    rc, res = gpu.verify_batch([(pk, sig, b"", msgs)])
    assert res == [(0, True)]


def test_staged_batch_relaunch_is_stable(gpu, oracle):
    rng = random.Random(5)
    (sk, pk), = _issuers(gpu, rng, 1)
    reqs = []
    for i in range(16):
        msgs = [rnd_bytes(rng, 32) for _ in range(rng.randrange(1, 9))]
        sig = oracle.sign(sk, pk, b"", msgs)
        if i % 4 == 3:
            msgs = msgs + [b"extra"]
        reqs.append((pk, sig, b"", msgs))
    b = gpu.StagedBatch(gpu.BATCH_VERIFY, reqs)
    first = None
    for _ in range(3):
        b.launch()
        b.sync()
        r = b.results()
        first = first or r
        assert r == first
    assert [v for _, v in first] == [i % 4 != 3 for i in range(16)]
    b.free()


def _on_curve_outside_g1(oracle):
    """Compressed encoding of a point on E(Fp) outside the prime-order subgroup."""
    p = 0x1A0111EA397FE69A4B1BA7B6434BACD764774B84F38512BF6730D2A0F6B0F6241EABFFFEB153FFFFB9FEFFFFFFFFAAAB
    for x in range(1, 400):
        rhs = (x ** 3 + 4) % p
        y = pow(rhs, (p + 1) // 4, p)
        if y * y % p == rhs:
            enc = bytearray(x.to_bytes(48, "big"))
            enc[0] |= 0x80
            assert not oracle.g1_valid(bytes(enc))
            return bytes(enc)
    raise AssertionError("no point found")


def test_points_outside_subgroup_rejected(gpu, oracle):
    """Verify's A and ProofVerify's Abar / Bbar / D on the curve but outside G1
    (the subgroup check runs beside the GLV products) give code 1; the other
    items of the batch are unaffected."""
    rng = random.Random(41)
    (sk, pk), = _issuers(gpu, rng, 1)
    bad = _on_curve_outside_g1(oracle)
    msgs = [rng.randbytes(40) for _ in range(6)]
    sig = oracle.sign(sk, pk, b"hdr", msgs)
    rc, res = gpu.verify_batch([(pk, sig, b"hdr", msgs), (pk, bad + sig[48:], b"hdr", msgs)])
    assert rc == 0 and res == [(0, True), (1, False)]
    idx = [1, 4]
    rc, pr = gpu.proof_gen_batch([(pk, sig, b"hdr", b"ph", msgs, idx)], seeds=[bytes(32)])
    assert rc == 0 and pr[0][0] == 0
    proof = pr[0][1]
    reqs = [(pk, proof, b"hdr", b"ph", [msgs[i] for i in idx], idx, len(msgs))]
    for k in range(3):  # Abar, Bbar, D
        broken = proof[:48 * k] + bad + proof[48 * (k + 1):]
        reqs.append((pk, broken, b"hdr", b"ph", [msgs[i] for i in idx], idx, len(msgs)))
    rc, res = gpu.proof_verify_batch(reqs)
    assert rc == 0
    assert res[0] == (0, True)
    assert [c for c, _ in res[1:]] == [1, 1, 1]


def test_zero_messages_all_operations(gpu, oracle):
    """L = 0: a signature and a proof over no messages (R = U = 0), with and
    without header / presentation header, against the oracle."""
    rng = random.Random(404)
    (sk, pk), = _issuers(gpu, rng, 1)
    cases = [(b"", b""), (b"hdr", b""), (b"", b"ph"), (rnd_bytes(rng, 64), rnd_bytes(rng, 64))]
    rc, sigs = gpu.sign_batch([(sk, pk, h, []) for h, _ in cases])
    assert rc == 0
    for (h, _), (code, sig) in zip(cases, sigs):
        assert (code, sig) == (0, oracle.sign(sk, pk, h, []))
    rc, res = gpu.verify_batch([(pk, sig, h, []) for (h, _), (_, sig) in zip(cases, sigs)] +
                               [(pk, sigs[0][1], b"x", [])])
    assert (rc, res) == (0, [(0, True)] * 4 + [(0, False)])
    seeds = [rnd_bytes(rng, 32) for _ in cases]
    greqs = [(pk, sig, h, ph, [], []) for (h, ph), (_, sig) in zip(cases, sigs)]
    rc, proofs = gpu.proof_gen_batch(greqs, seeds=seeds)
    assert rc == 0
    for (h, ph), (_, sig), seed, (code, proof) in zip(cases, sigs, seeds, proofs):
        assert code == 0 and len(proof) == 272
        assert proof == oracle.proof_gen(pk, sig, h, ph, [], [], oracle.chacha_random_scalars(seed, 5))
    pv = [(pk, proof, h, ph, [], [], 0) for (h, ph), (_, proof) in zip(cases, proofs)]
    rc, res = gpu.proof_verify_batch(pv + [(pk, proofs[0][1], b"", b"other", [], [], 0)])
    assert (rc, res) == (0, [(0, True)] * 4 + [(0, False)])
