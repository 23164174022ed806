"""The bbs-2023 proofValue codec (crates/cubbs-bbs-proofvalue/src/lib.rs).

The oracle (oracle/proofvalue_oracle.py) is pinned first: RFC 8949 Appendix A
(CBOR heads and the definite/indefinite decoder), RFC 4648 section 10 and
Python's base64 module (base64url), and the reference's own two tests
(lib.rs:191-220).  The native codec (include/cubbs_proofvalue.h) is then held
to the oracle on seeded random values, on hand-built alternative encodings,
and on a corpus of malformed inputs where the error variant must agree.
All host code: no GPU needed."""
import base64
import os
import random
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cubbs-cred_amd"))
import proofvalue_oracle as O  # noqa: E402
import proofvalue as PV  # noqa: E402


# ---------------------------------------------------------------- oracle pins
RFC4648 = [(b"", ""), (b"f", "Zg"), (b"fo", "Zm8"), (b"foo", "Zm9v"), (b"foob", "Zm9vYg"),
           (b"fooba", "Zm9vYmE"), (b"foobar", "Zm9vYmFy")]


def test_oracle_base64url_rfc4648_and_library_routine():
    for raw, text in RFC4648:
        assert O.b64url_encode(raw) == text
        assert O.b64url_decode(text) == raw
    rng = random.Random(1)
    for n in range(0, 70):
        raw = bytes(rng.randrange(256) for _ in range(n))
        ref = base64.urlsafe_b64encode(raw).rstrip(b"=").decode()
        assert O.b64url_encode(raw) == ref
        assert O.b64url_decode(ref) == raw
    # no padding, no lone symbol, no spare bits, url-safe alphabet only
    for bad in ("Zg==", "Z", "Zh", "Zm9", "Zm+v", "Zm/v", "Zm9v\n", "Zm 9v"):
        with pytest.raises(O.ProofValueError) as e:
            O.b64url_decode(bad)
        assert e.value.kind == O.BASE64
    assert O.b64url_decode("Zm8") == b"fo" and O.b64url_decode("-_8") == bytes([0xFB, 0xFF])


RFC8949_HEADS = [(0, "00"), (1, "01"), (10, "0a"), (23, "17"), (24, "1818"), (25, "1819"), (100, "1864"),
                 (1000, "1903e8"), (1000000, "1a000f4240"), (1000000000000, "1b000000e8d4a51000"),
                 (18446744073709551615, "1bffffffffffffffff")]


def test_oracle_cbor_encoding_rfc8949():
    for v, hx in RFC8949_HEADS:
        assert O.enc_uint(v).hex() == hx
        assert O.cbor_decode_tree(bytes.fromhex(hx)) == ("uint", v)
    assert O.enc_bytes(b"").hex() == "40"
    assert O.enc_bytes(bytes([1, 2, 3, 4])).hex() == "4401020304"
    for s, hx in (("", "60"), ("a", "6161"), ("IETF", "6449455446"), ('"\\', "62225c"),
                  ("ü", "62c3bc"), ("水", "63e6b0b4")):
        assert O.enc_text(s).hex() == hx
    assert O.cbor_head(4, 0).hex() == "80" and O.cbor_head(5, 0).hex() == "a0"
    assert (O.cbor_head(4, 3) + b"".join(O.enc_uint(i) for i in (1, 2, 3))).hex() == "83010203"
    assert (O.cbor_head(5, 2) + bytes([1, 2, 3, 4])).hex() == "a201020304"


def test_oracle_cbor_decoding_rfc8949_indefinite_and_tags():
    T = O.cbor_decode_tree
    u = lambda v: ("uint", v)  # noqa: E731
    assert T(bytes.fromhex("5f42010243030405ff")) == ("bytes", bytes([1, 2, 3, 4, 5]))
    assert T(bytes.fromhex("7f657374726561646d696e67ff")) == ("text", b"streaming")
    assert T(bytes.fromhex("9fff")) == ("array", [])
    nested = ("array", [u(1), ("array", [u(2), u(3)]), ("array", [u(4), u(5)])])
    for hx in ("9f018202039f0405ffff", "9f01820203820405ff", "83018202039f0405ff", "83019f0203ff820405"):
        assert T(bytes.fromhex(hx)) == nested
    assert T(bytes.fromhex("9f0102030405060708090a0b0c0d0e0f101112131415161718181819ff")) == \
        ("array", [u(i) for i in range(1, 26)])
    assert T(bytes.fromhex("bf61610161629f0203ffff")) == \
        ("map", [(("text", b"a"), u(1)), (("text", b"b"), ("array", [u(2), u(3)]))])
    assert T(bytes.fromhex("d74401020304")) == ("tag", 23, ("bytes", bytes([1, 2, 3, 4])))
    assert T(bytes.fromhex("d818456449455446")) == ("tag", 24, ("bytes", b"dIETF"))
    assert T(bytes.fromhex("c11a514b67b0")) == ("tag", 1, u(1363896240))
    assert T(bytes.fromhex("3903e7")) == ("neg", -1000)
    for bad in ("18", "5f41ff", "5f6161ff", "62c3", "ff", "1c", "0000", "9f01", "a1"):
        with pytest.raises(O._CborFail):
            T(bytes.fromhex(bad))


def reference_base():
    # lib.rs:195-206
    return dict(sig=bytes([1]) * 80, header=bytes([2]) * 64, pk=bytes([3]) * 96, hmac=bytes([4]) * 32,
                pointers=["/issuer"])


def reference_derived():
    # lib.rs:208-219
    return dict(proof=bytes([1]) * 272, labels={"c14n0": "b2", "c14n1": "b0"}, mandatory=[0, 1, 4],
                selective=[2, 3], ph=b"ph")


def test_oracle_reference_round_trips():
    b = reference_base()
    pv = O.serialize_base(b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"])
    assert pv.startswith("u2V0C")
    assert O.parse_base(pv) == (b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"])
    d = reference_derived()
    pv = O.serialize_derived(d["proof"], d["labels"], d["mandatory"], d["selective"], d["ph"])
    assert pv.startswith("u2V0D")
    assert O.parse_derived(pv) == (d["proof"], d["labels"], d["mandatory"], d["selective"], d["ph"])


# ---------------------------------------------------------------- library vs oracle
def lib_base(b):
    return PV.serialize_base(PV.BaseProofValue(b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"]))


def lib_derived(d):
    return PV.serialize_derived(PV.DerivedProofValue(d["proof"], d["labels"], d["mandatory"], d["selective"],
                                                     d["ph"]))


def lib_parse_base(pv):
    v = PV.parse_base(pv)
    return (v.bbs_signature, v.bbs_header, v.public_key, v.hmac_key, v.mandatory_pointers)


def lib_parse_derived(pv):
    v = PV.parse_derived(pv)
    return (v.bbs_proof, v.label_map, v.mandatory_indexes, v.selective_indexes, v.presentation_header)


def outcome(fn, *args):
    try:
        return ("ok", fn(*args))
    except (O.ProofValueError, PV.ProofValueError) as e:
        return ("err", e.kind)


def test_library_exports_codec_symbols():
    lib = PV.lib()
    for name in PV.SYMBOLS:
        assert hasattr(lib, name)


def test_library_reference_round_trips():
    b = reference_base()
    pv = lib_base(b)
    assert pv.startswith("u2V0C")
    assert pv == O.serialize_base(b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"])
    assert lib_parse_base(pv) == (b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"])
    d = reference_derived()
    pv = lib_derived(d)
    assert pv.startswith("u2V0D")
    assert pv == O.serialize_derived(d["proof"], d["labels"], d["mandatory"], d["selective"], d["ph"])
    assert lib_parse_derived(pv) == (d["proof"], d["labels"], d["mandatory"], d["selective"], d["ph"])


def rand_u64(rng):
    return rng.choice([0, 1, 23, 24, 255, 256, 65535, 65536, 2**32 - 1, 2**32, 2**64 - 1, rng.randrange(2**64),
                       rng.randrange(1000)])


def rand_text(rng):
    alphabet = "/abcXYZ~01_-ü水\U0001f600"
    return "".join(rng.choice(alphabet) for _ in range(rng.randrange(0, 30)))


def rand_base(rng):
    return dict(sig=rng.randbytes(80), header=rng.randbytes(64), pk=rng.randbytes(96), hmac=rng.randbytes(32),
                pointers=[rand_text(rng) for _ in range(rng.choice([0, 1, 2, 5, 30]))])


def rand_derived(rng):
    labels = {}
    for _ in range(rng.choice([0, 1, 3, 40])):
        labels["c14n%d" % rand_u64(rng)] = "b%d" % rand_u64(rng)
    return dict(proof=rng.randbytes(rng.choice([1, 23, 24, 272, 1040, 70000])), labels=labels,
                mandatory=[rand_u64(rng) for _ in range(rng.choice([0, 2, 30]))],
                selective=[rand_u64(rng) for _ in range(rng.choice([0, 2, 24, 300]))],
                ph=rng.randbytes(rng.choice([0, 2, 64, 300])))


def test_library_matches_oracle_on_random_values():
    rng = random.Random(7)
    for _ in range(150):
        b = rand_base(rng)
        pv = lib_base(b)
        assert pv == O.serialize_base(b["sig"], b["header"], b["pk"], b["hmac"], b["pointers"])
        assert lib_parse_base(pv) == O.parse_base(pv) == (b["sig"], b["header"], b["pk"], b["hmac"],
                                                            b["pointers"])
        d = rand_derived(rng)
        pv = lib_derived(d)
        assert pv == O.serialize_derived(d["proof"], d["labels"], d["mandatory"], d["selective"], d["ph"])
        got = lib_parse_derived(pv)
        assert got == O.parse_derived(pv)
        assert got[0] == d["proof"] and got[2:] == (d["mandatory"], d["selective"], d["ph"])


# hand-built CBOR: ("h", major, value, width) writes a head of a chosen width
def enc(tree):
    kind = tree[0]
    if kind == "raw":
        return tree[1]
    if kind == "h":
        _, major, v, width = tree
        ai = {0: None, 1: 24, 2: 25, 4: 26, 8: 27}[width]
        return bytes([major << 5 | v]) if ai is None else bytes([major << 5 | ai]) + v.to_bytes(width, "big")
    if kind == "uint":
        return O.enc_uint(tree[1])
    if kind == "bytes":
        return O.enc_bytes(tree[1])
    if kind == "text":
        return O.cbor_head(3, len(tree[1])) + tree[1]
    if kind == "ibytes":   # indefinite byte string from chunks
        return b"\x5f" + b"".join(O.enc_bytes(c) for c in tree[1]) + b"\xff"
    if kind == "itext":
        return b"\x7f" + b"".join(O.cbor_head(3, len(c)) + c for c in tree[1]) + b"\xff"
    if kind == "array":
        return O.cbor_head(4, len(tree[1])) + b"".join(enc(x) for x in tree[1])
    if kind == "iarray":
        return b"\x9f" + b"".join(enc(x) for x in tree[1]) + b"\xff"
    if kind == "map":
        return O.cbor_head(5, len(tree[1])) + b"".join(enc(k) + enc(v) for k, v in tree[1])
    if kind == "imap":
        return b"\xbf" + b"".join(enc(k) + enc(v) for k, v in tree[1]) + b"\xff"
    if kind == "tag":
        return O.cbor_head(6, tree[1]) + enc(tree[2])
    raise ValueError(kind)


def wrap(prefix, body):
    return "u" + O.b64url_encode(prefix + body)


def base_tree(sig=None, header=None, pk=None, hmac=None, pointers=None):
    return ["array", [sig or ("bytes", bytes(80)), header or ("bytes", bytes(64)), pk or ("bytes", bytes(96)),
                      hmac or ("bytes", bytes(32)), pointers or ("array", [("text", b"/a")])]]


def derived_tree(proof=None, labels=None, man=None, sel=None, ph=None):
    return ["array", [proof or ("bytes", b"\x07" * 5), labels or ("map", [(("uint", 3), ("uint", 1))]),
                      man or ("array", [("uint", 0)]), sel or ("array", [("uint", 2)]), ph or ("bytes", b"")]]


def both(kind, pv):
    if kind == "base":
        a, b = outcome(lib_parse_base, pv), outcome(O.parse_base, pv)
    else:
        a, b = outcome(lib_parse_derived, pv), outcome(O.parse_derived, pv)
    assert a == b, (pv, a, b)
    return a


def test_alternative_encodings_accepted_alike():
    B, D = O.PREFIX_BASE, O.PREFIX_DERIVED
    cases_base = [
        base_tree(sig=("ibytes", [bytes(40), b"", bytes(40)])),
        base_tree(sig=("text", b"a" * 80)),                                  # text in a bytes field
        base_tree(hmac=("array", [("uint", i) for i in range(32)])),          # sequence of u8
        base_tree(hmac=("iarray", [("h", 0, i, 8) for i in range(32)])),     # non-minimal u8
        base_tree(pointers=("array", [("bytes", b"/x"), ("itext", [b"/", b"y"])])),  # bytes in a text field
        base_tree(pointers=("iarray", [])),
        ["iarray", base_tree()[1]],
        ("tag", 55799, base_tree()),                                          # self-describe tag
        base_tree(pk=("tag", 24, ("tag", 2, ("bytes", bytes(96))))),
        base_tree(sig=("itext", [b"\xc3", b"\xbc" + b"a" * 78])),             # utf-8 checked on the whole
    ]
    for t in cases_base:
        r = both("base", wrap(B, enc(t)))
        assert r[0] == "ok", (t, r)
    cases_derived = [
        derived_tree(labels=("imap", [(("uint", 5), ("uint", 6)), (("uint", 5), ("uint", 9))])),  # dup: last wins
        derived_tree(labels=("map", [(("h", 0, 10, 8), ("h", 0, 1, 2))])),
        derived_tree(man=("iarray", [("h", 0, 7, 4), ("tag", 1, ("uint", 2**64 - 1))])),
        derived_tree(proof=("array", [("uint", 255)])),
        derived_tree(ph=("itext", [b"ph"])),
        ("tag", 0, ["iarray", derived_tree()[1]]),
    ]
    for t in cases_derived:
        r = both("derived", wrap(D, enc(t)))
        assert r[0] == "ok", (t, r)
    r = both("derived", wrap(D, enc(cases_derived[0])))
    assert r[1][1] == {"c14n5": "b9"}


def test_malformed_inputs_rejected_alike():
    B, D = O.PREFIX_BASE, O.PREFIX_DERIVED
    expect = [
        ("base", "", "InvalidMultibasePrefix"),
        ("base", "z2V0C", "InvalidMultibasePrefix"),
        ("base", "u2V0C=", "Base64"),
        ("base", "u2V0Cg", "Base64"),                     # lone final symbol
        ("base", "u2V1", "Base64"),                       # spare bits set
        ("base", "u", "InvalidCborPrefix"),
        ("base", wrap(D, enc(base_tree())), "InvalidCborPrefix"),
        ("derived", wrap(B, enc(derived_tree())), "InvalidCborPrefix"),
        ("base", wrap(B, b""), "Cbor"),
        ("base", wrap(B, enc(base_tree()) + b"\x00"), "Cbor"),                    # trailing data
        ("base", wrap(B, enc(["array", base_tree()[1][:4]])), "Cbor"),            # four fields
        ("base", wrap(B, enc(["array", base_tree()[1] + [("uint", 1)]])), "Cbor"),  # six fields
        ("base", wrap(B, enc(["iarray", base_tree()[1] + [("uint", 1)]])), "Cbor"),
        ("base", wrap(B, enc(base_tree(sig=("uint", 5)))), "Cbor"),
        ("base", wrap(B, enc(base_tree(hmac=("array", [("uint", 256)] * 32)))), "Cbor"),
        ("base", wrap(B, enc(base_tree(pointers=("array", [("bytes", b"\xff")])))), "Cbor"),
        ("base", wrap(B, enc(base_tree(pointers=("array", [("text", b"\xed\xa0\x80")])))), "Cbor"),
        ("base", wrap(B, enc(base_tree(sig=("raw", b"\x5f\x41\x00\x60\xff")))), "Cbor"),  # mixed chunks
        ("base", wrap(B, enc(base_tree(sig=("raw", b"\x5f\x5f\xff\xff")))), "Cbor"),      # nested indefinite
        ("base", wrap(B, enc(base_tree(sig=("raw", b"\xfb" + bytes(8))))), "Cbor"),
        ("base", wrap(B, enc(base_tree(sig=("raw", b"\xf6")))), "Cbor"),
        ("base", wrap(B, enc(base_tree(sig=("raw", b"\xdc\x40")))), "Cbor"),               # unassigned tag head
        ("base", wrap(B, enc(base_tree(sig=("bytes", bytes(79))))), "InvalidBaseProof"),
        ("base", wrap(B, enc(base_tree(hmac=("bytes", bytes(33))))), "InvalidBaseProof"),
        ("derived", wrap(D, enc(derived_tree(proof=("bytes", b"")))), "InvalidDerivedProof"),
        ("derived", wrap(D, enc(derived_tree(labels=("map", [(("text", b"1"), ("uint", 1))])))), "Cbor"),
        ("derived", wrap(D, enc(derived_tree(man=("array", [("raw", b"\x20")])))), "Cbor"),  # negative
        ("derived", wrap(D, enc(derived_tree(man=("array", [("raw", b"\x1c")])))), "Cbor"),
        ("derived", wrap(D, enc(derived_tree())[:-1]), "Cbor"),
        ("derived", wrap(D, enc(derived_tree(labels=("raw", b"\xbf\x01\x01")))), "Cbor"),
    ]
    for kind, pv, want in expect:
        r = both(kind, pv)
        assert r == ("err", want), (pv, r, want)


def test_truncations_and_mutations_agree():
    rng = random.Random(11)
    samples = []
    for _ in range(4):
        b = rand_base(rng)
        b["pointers"] = b["pointers"][:3]
        samples.append(("base", O.PREFIX_BASE,
                        O.b64url_decode(O.serialize_base(b["sig"], b["header"], b["pk"], b["hmac"],
                                                         b["pointers"])[1:])))
        d = rand_derived(rng)
        d["proof"] = d["proof"][:40]
        d["selective"] = d["selective"][:5]
        samples.append(("derived", O.PREFIX_DERIVED,
                        O.b64url_decode(O.serialize_derived(d["proof"], d["labels"], d["mandatory"],
                                                            d["selective"], d["ph"])[1:])))
    seen = set()
    for kind, prefix, raw in samples:
        for cut in range(len(raw)):
            seen.add(both(kind, "u" + O.b64url_encode(raw[:cut]))[1])
        for _ in range(400):
            m = bytearray(raw)
            for _ in range(rng.choice([1, 1, 2, 3])):
                j = rng.randrange(3, len(m))
                m[j] = rng.randrange(256) if rng.random() < 0.5 else m[j] ^ (1 << rng.randrange(8))
            r = both(kind, "u" + O.b64url_encode(bytes(m)))
            seen.add(r[1] if r[0] == "err" else "ok")
    assert {"Cbor", "InvalidCborPrefix", "ok"} <= seen


def test_nesting_limit_agrees():
    # top array + list array + n tags on one index: serde_cbor's 128th level fails
    D = O.PREFIX_DERIVED
    for n, ok in ((125, True), (126, False)):
        t = ("uint", 1)
        for _ in range(n):
            t = ("tag", 9, t)
        r = both("derived", wrap(D, enc(derived_tree(man=("array", [t])))))
        assert (r[0] == "ok") == ok, (n, r)


def test_serialize_label_map_rules():
    base = dict(proof=b"\x01", mandatory=[], selective=[], ph=b"")
    cases = [
        ({"c14n+7": "b+2"}, {"c14n7": "b2"}),
        ({"c14n007": "b0"}, {"c14n7": "b0"}),
        ({"c14n18446744073709551615": "b18446744073709551615"},
         {"c14n18446744073709551615": "b18446744073709551615"}),
        ({"c14n01": "b5", "c14n1": "b6"}, {"c14n1": "b6"}),   # key order: "c14n01" < "c14n1", last wins
        ({"c14n10": "b1", "c14n9": "b2"}, {"c14n10": "b1", "c14n9": "b2"}),
    ]
    for labels, want in cases:
        d = dict(base, labels=labels)
        pv = lib_derived(d)
        assert pv == O.serialize_derived(d["proof"], labels, [], [], b"")
        assert lib_parse_derived(pv)[1] == want
        assert list(PV.parse_derived(pv).label_map) == sorted(want)
    for labels in ({"c14n": "b1"}, {"c14nx": "b1"}, {"c14n-1": "b1"}, {"C14n1": "b1"}, {"c14n+": "b1"},
                   {"c14n18446744073709551616": "b1"}, {"c14n1": "b"}, {"c14n1": "1"}, {"c14n1": "b 1"},
                   {"c14n1": "b-1"}):
        d = dict(base, labels=labels)
        assert outcome(lib_derived, d) == outcome(O.serialize_derived, d["proof"], labels, [], [], b"") == \
            ("err", "InvalidDerivedProof"), labels
    d = dict(base, labels={}, proof=b"")
    assert outcome(lib_derived, d) == ("err", "InvalidDerivedProof")
    b = reference_base()
    for key, n in (("sig", 79), ("header", 65), ("pk", 0), ("hmac", 31)):
        bad = dict(b, **{key: bytes(n)})
        assert outcome(lib_base, bad) == ("err", "InvalidBaseProof")
        assert outcome(O.serialize_base, bad["sig"], bad["header"], bad["pk"], bad["hmac"], []) == \
            ("err", "InvalidBaseProof")


def test_c_abi_buffers_and_arguments():
    import ctypes
    lib = PV.lib()
    b = reference_base()
    pv = lib_base(b).encode()
    need = ctypes.c_size_t(0)
    out = PV.CBase()
    assert lib.cubbs_proofvalue_parse_base(pv, len(pv), None, 0, ctypes.byref(need), ctypes.byref(out)) == \
        PV.ERR_INSUFFICIENT_OUTPUT
    assert need.value >= 80 + 64 + 96 + 32
    # a misaligned arena one byte short fails, of the exact size succeeds
    raw = ctypes.create_string_buffer(need.value + 1)
    arena = ctypes.cast(ctypes.addressof(raw) + 1, ctypes.c_void_p)
    assert lib.cubbs_proofvalue_parse_base(pv, len(pv), arena, need.value - 1, ctypes.byref(need),
                                           ctypes.byref(out)) == PV.ERR_INSUFFICIENT_OUTPUT
    assert lib.cubbs_proofvalue_parse_base(pv, len(pv), arena, need.value, ctypes.byref(need),
                                           ctypes.byref(out)) == PV.OK
    assert ctypes.string_at(out.public_key.ptr, out.public_key.len) == b["pk"]
    assert out.mandatory_pointers[0].len == len("/issuer")
    # decode errors come before buffer sizes, and leave the size 0
    bad = b"u2V0@"
    assert lib.cubbs_proofvalue_parse_base(bad, len(bad), None, 0, ctypes.byref(need), ctypes.byref(out)) == \
        PV.ERR_BASE64 and need.value == 0
    # a text field must be UTF-8 (a Rust String is)
    with pytest.raises(PV.ProofValueError) as e:
        PV.serialize_base(PV.BaseProofValue(b["sig"], b["header"], b["pk"], b["hmac"], ["\udc80"]))
    assert e.value.kind == "Argument"
    # serialized length is reported when the buffer is short
    c = PV.CBase()
    keep = PV._Keep()
    c.bbs_signature, c.bbs_header = keep.slice(b["sig"]), keep.slice(b["header"])
    c.public_key, c.hmac_key = keep.slice(b["pk"]), keep.slice(b["hmac"])
    small = ctypes.create_string_buffer(10)
    assert lib.cubbs_proofvalue_serialize_base(ctypes.byref(c), small, 10, ctypes.byref(need)) == \
        PV.ERR_INSUFFICIENT_OUTPUT
    payload = 3 + 1 + (2 + 80) + (2 + 64) + (2 + 96) + (2 + 32) + 1
    assert need.value == 1 + len(O.b64url_encode(bytes(payload)))


def test_codec_under_sanitizers(tmp_path):
    """The codec's C entries over 60k mutated inputs, built with g++ under
    AddressSanitizer and UBSan (tests/native/proofvalue_fuzz.cpp): no
    report, and both decode errors and successes occur."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "pv_fuzz")
    subprocess.check_call(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                           "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "native", "proofvalue_fuzz.cpp"),
                           os.path.join(ROOT, "cubbs-cred_amd", "csrc", "proofvalue.cpp"), "-o", exe])
    out = subprocess.run([exe, "60000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-3000:]
    counts = [int(x) for x in out.stdout.split()]
    assert len(counts) == 10 and sum(counts) == 60000
    assert counts[0] > 0 and counts[6] > 0   # successes and CBOR errors both occur
