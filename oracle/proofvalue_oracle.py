"""Oracle for the bbs-2023 proofValue codec -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module; the product codec is the C++ one in
cubbs-cred_amd/csrc/proofvalue.cpp.  This is a plain-Python restatement of
crates/cubbs-bbs-proofvalue/src/lib.rs and of the published behaviour of the
crates it builds on (versions pinned in the reference's Cargo.lock):
serde_cbor 0.11.2 (decoding and encoding), serde_bytes 0.11.19 (the ByteBuf
visitor) and base64 0.22.1 (URL_SAFE_NO_PAD).  Unlike the product's
streaming reader, it decodes CBOR into a tree first and then applies serde's
visitor rules to the tree.

Pinned by tests/test_proofvalue.py against: the RFC 8949 Appendix A
examples (encoder heads, decoder of definite and indefinite forms), the
RFC 4648 section 10 base64 vectors and Python's base64 module, and the
reference's own tests (lib.rs:191-220: prefixes "u2V0C" / "u2V0D" and round
trips).  The decoder's acceptance of serde-specific alternatives (text or an
array of u8 in a bytes field, bytes in a text field, tags skipped, the depth
limit) is restated from those crates' sources as published and has no golden
vector in the reference: parity unpinned for those inputs beyond agreement
of the two independent implementations.
"""

PREFIX_BASE = bytes([0xD9, 0x5D, 0x02])      # lib.rs:11
PREFIX_DERIVED = bytes([0xD9, 0x5D, 0x03])   # lib.rs:12

# Error variants (lib.rs:16-35), named as in the reference.
MULTIBASE, CBOR_PREFIX, INVALID_BASE, INVALID_DERIVED, BASE64, CBOR = (
    "InvalidMultibasePrefix", "InvalidCborPrefix", "InvalidBaseProof",
    "InvalidDerivedProof", "Base64", "Cbor")


class ProofValueError(Exception):
    def __init__(self, kind):
        super().__init__(kind)
        self.kind = kind


# ---------------------------------------------------------------- base64url no pad
ALPHABET = b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789-_"


def b64url_encode(data):
    out = []
    for i in range(0, len(data), 3):
        chunk = data[i:i + 3]
        n = int.from_bytes(chunk + bytes(3 - len(chunk)), "big")
        syms = [(n >> s) & 63 for s in (18, 12, 6, 0)]
        out.extend(ALPHABET[x] for x in syms[:len(chunk) + 1])
    return bytes(out).decode()


def b64url_decode(text):
    """base64 0.22 URL_SAFE_NO_PAD: alphabet only, no padding, no lone final
    symbol, unused bits of the last symbol zero."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    if len(raw) % 4 == 1:
        raise ProofValueError(BASE64)
    vals = []
    for c in raw:
        k = ALPHABET.find(bytes([c]))
        if k < 0:
            raise ProofValueError(BASE64)
        vals.append(k)
    out = bytearray()
    for i in range(0, len(vals), 4):
        group = vals[i:i + 4]
        n = 0
        for v in group:
            n = n * 64 + v
        nbytes = len(group) * 6 // 8
        spare = len(group) * 6 - nbytes * 8
        if n & ((1 << spare) - 1):
            raise ProofValueError(BASE64)
        out += (n >> spare).to_bytes(nbytes, "big")
    return bytes(out)


# ---------------------------------------------------------------- CBOR encoding (serde_cbor)
def cbor_head(major, value):
    """Shortest head, as serde_cbor's write_u64 picks it."""
    m = major << 5
    if value < 24:
        return bytes([m | value])
    for ai, width in ((24, 1), (25, 2), (26, 4), (27, 8)):
        if value < 1 << (8 * width):
            return bytes([m | ai]) + value.to_bytes(width, "big")
    raise ValueError("value exceeds 64 bits")


def enc_bytes(b):
    return cbor_head(2, len(b)) + bytes(b)


def enc_text(s):
    b = s.encode()
    return cbor_head(3, len(b)) + b


def enc_uint(v):
    return cbor_head(0, v)


# ---------------------------------------------------------------- CBOR decoding to a tree
class _CborFail(Exception):
    pass


def utf8_valid(b):
    try:
        bytes(b).decode("utf-8", errors="strict")
        return True
    except UnicodeDecodeError:
        return False


def cbor_decode_tree(buf):
    """Decodes one CBOR item that must span buf exactly.  Items are tuples:
    ("uint", v) ("neg", v) ("bytes", b) ("text", b) ("array", [..])
    ("map", [(k, v)..]) ("tag", t, item) ("simple", code) ("float", raw).
    Raises _CborFail on malformed input.  Arrays, maps and tags each use one
    of serde_cbor's 128 nesting levels (the 128th fails)."""
    pos = [0]
    depth = [128]

    def need(k):
        if pos[0] + k > len(buf):
            raise _CborFail("eof")
        r = buf[pos[0]:pos[0] + k]
        pos[0] += k
        return r

    def argument(ai):
        if ai < 24:
            return ai
        if ai <= 27:
            return int.from_bytes(need(1 << (ai - 24)), "big")
        raise _CborFail("reserved additional information")

    def nest():
        depth[0] -= 1
        if depth[0] == 0:
            raise _CborFail("recursion limit")

    def item():
        ib = need(1)[0]
        major, ai = ib >> 5, ib & 31
        if ib == 0xFF:
            raise _CborFail("unexpected break")
        if major in (0, 1):
            v = argument(ai)
            return ("uint", v) if major == 0 else ("neg", -1 - v)
        if major in (2, 3):
            if ai == 31:
                data = bytearray()
                while True:
                    c = need(1)[0]
                    if c == 0xFF:
                        break
                    if c >> 5 != major or c & 31 == 31:
                        raise _CborFail("bad chunk")
                    data += need(argument(c & 31))
                data = bytes(data)
            else:
                data = bytes(need(argument(ai)))
            if major == 3 and not utf8_valid(data):
                raise _CborFail("invalid utf-8")
            return ("bytes" if major == 2 else "text", data)
        if major in (4, 5):
            nest()
            entries = []

            def one():
                if major == 4:
                    entries.append(item())
                else:
                    k = item()
                    entries.append((k, item()))
            if ai == 31:
                while True:
                    if pos[0] >= len(buf):
                        raise _CborFail("eof")
                    if buf[pos[0]] == 0xFF:
                        pos[0] += 1
                        break
                    one()
            else:
                for _ in range(argument(ai)):
                    one()
            depth[0] += 1
            return ("array" if major == 4 else "map", entries)
        if major == 6:
            tag = argument(ai)
            nest()
            inner = item()
            depth[0] += 1
            return ("tag", tag, inner)
        # major 7: simple values and floats; none fits a payload field
        if ai < 24:
            return ("simple", ai)
        if ai == 24:
            return ("simple", need(1)[0])
        if ai in (25, 26, 27):
            return ("float", need(1 << (ai - 24)))
        raise _CborFail("reserved simple")

    top = item()
    if pos[0] != len(buf):
        raise _CborFail("trailing data")
    return top


# ---------------------------------------------------------------- serde visitors over the tree
def _untag(it):
    while it[0] == "tag":
        it = it[2]
    return it


def as_u64(it):
    it = _untag(it)
    if it[0] != "uint":
        raise _CborFail("not an unsigned integer")
    return it[1]


def as_bytes(it):
    """serde_bytes ByteBuf: bytes, text, or a sequence of u8."""
    it = _untag(it)
    if it[0] in ("bytes", "text"):
        return it[1]
    if it[0] == "array":
        vals = [as_u64(x) for x in it[1]]
        if any(v > 255 for v in vals):
            raise _CborFail("not a u8")
        return bytes(vals)
    raise _CborFail("not bytes")


def as_text(it):
    """serde String: text, or bytes that are valid UTF-8."""
    it = _untag(it)
    if it[0] == "text" or (it[0] == "bytes" and utf8_valid(it[1])):
        return it[1].decode()
    raise _CborFail("not text")


def as_seq(it):
    it = _untag(it)
    if it[0] != "array":
        raise _CborFail("not a sequence")
    return it[1]


def as_u64_map(it):
    it = _untag(it)
    if it[0] != "map":
        raise _CborFail("not a map")
    out = {}
    for k, v in it[1]:
        out[as_u64(k)] = as_u64(v)  # later duplicates win (BTreeMap insert)
    return out


def as_tuple5(it):
    fields = as_seq(it)
    if len(fields) != 5:
        raise _CborFail("payload arity")
    return fields


# ---------------------------------------------------------------- lib.rs restated
def _u64_from_str(s):
    """Rust's u64::from_str: optional '+', one or more ASCII digits, < 2^64."""
    if s.startswith("+"):
        s = s[1:]
    if not s or any(c not in "0123456789" for c in s):
        return None
    v = int(s)
    return v if v < 1 << 64 else None


def serialize_base(sig, header, pk, hmac_key, pointers):
    """lib.rs:73-83 and validate_base lib.rs:145-153."""
    if (len(sig), len(header), len(pk), len(hmac_key)) != (80, 64, 96, 32):
        raise ProofValueError(INVALID_BASE)
    body = cbor_head(4, 5) + enc_bytes(sig) + enc_bytes(header) + enc_bytes(pk) + enc_bytes(hmac_key)
    body += cbor_head(4, len(pointers)) + b"".join(enc_text(p) for p in pointers)
    return "u" + b64url_encode(PREFIX_BASE + body)


def compress_label_map(labels):
    """lib.rs:165-183; labels is a dict read in key order."""
    out = {}
    for k in sorted(labels, key=lambda s: s.encode()):
        v = labels[k]
        a = _u64_from_str(k[4:]) if k.startswith("c14n") else None
        b = _u64_from_str(v[1:]) if v.startswith("b") else None
        if a is None or b is None:
            raise ProofValueError(INVALID_DERIVED)
        out[a] = b
    return out


def serialize_derived(proof, labels, mandatory, selective, ph):
    """lib.rs:98-108: validate_derived, compress, encode."""
    if not proof:
        raise ProofValueError(INVALID_DERIVED)
    comp = compress_label_map(labels)
    body = cbor_head(4, 5) + enc_bytes(proof) + cbor_head(5, len(comp))
    for k in sorted(comp):
        body += enc_uint(k) + enc_uint(comp[k])
    body += cbor_head(4, len(mandatory)) + b"".join(enc_uint(x) for x in mandatory)
    body += cbor_head(4, len(selective)) + b"".join(enc_uint(x) for x in selective)
    body += enc_bytes(ph)
    return "u" + b64url_encode(PREFIX_DERIVED + body)


def _unwrap(proof_value, prefix):
    """lib.rs:132-143."""
    if not proof_value.startswith("u"):
        raise ProofValueError(MULTIBASE)
    raw = b64url_decode(proof_value[1:])
    if raw[:3] != prefix:
        raise ProofValueError(CBOR_PREFIX)
    try:
        return cbor_decode_tree(raw[3:])
    except _CborFail:
        raise ProofValueError(CBOR)


def parse_base(proof_value):
    """lib.rs:85-96 -> (sig, header, pk, hmac_key, [pointers])."""
    tree = _unwrap(proof_value, PREFIX_BASE)
    try:
        f = as_tuple5(tree)
        value = (as_bytes(f[0]), as_bytes(f[1]), as_bytes(f[2]), as_bytes(f[3]),
                 [as_text(x) for x in as_seq(f[4])])
    except _CborFail:
        raise ProofValueError(CBOR)
    if tuple(len(x) for x in value[:4]) != (80, 64, 96, 32):
        raise ProofValueError(INVALID_BASE)
    return value


def parse_derived(proof_value):
    """lib.rs:110-121 -> (proof, {c14n label: b label}, [mandatory],
    [selective], ph); decompress_label_map lib.rs:185-189."""
    tree = _unwrap(proof_value, PREFIX_DERIVED)
    try:
        f = as_tuple5(tree)
        proof = as_bytes(f[0])
        comp = as_u64_map(f[1])
        mandatory = [as_u64(x) for x in as_seq(f[2])]
        selective = [as_u64(x) for x in as_seq(f[3])]
        ph = as_bytes(f[4])
    except _CborFail:
        raise ProofValueError(CBOR)
    if not proof:
        raise ProofValueError(INVALID_DERIVED)
    labels = {"c14n%d" % k: "b%d" % v for k, v in comp.items()}
    return proof, labels, mandatory, selective, ph
