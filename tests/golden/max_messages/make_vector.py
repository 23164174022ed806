"""Writes l65536.json: the oracle's Sign output, and its seeded ProofGen
output, for one signature over 65,536 messages.  This is the maximum-size
case of the reference's engine test
crates/cubbs-bbs/src/tests.rs:894-1030
(loaded_gpu_65536_four_operations_match_cpu_and_reject_tampering): the same
64-byte header and message construction, with a seeded key in place of the
reference's thread_rng key.  The proof uses ProofGen's seeded-ChaCha20
blinding convention and discloses the even indexes; being 1 MB, it is kept
as its length and SHA-256.

Run from the repository root after building oracle/ (takes minutes: the
oracle creates 65,537 generators and evaluates the MSMs term by term)."""
import hashlib
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "..", "oracle"))
import oracle_py as o  # noqa: E402

L = 65536
IKM = bytes(range(32))
SEED = bytes(range(100, 132))
PH = b"omniid-route4-65536-presentation"


def header():
    h = bytearray(64)
    h[:8] = L.to_bytes(8, "big")
    h[8:33] = b"omniid-route4-correctness"
    return bytes(h)


def messages():
    return [i.to_bytes(8, "big") + b"route4-message-fixture!!" for i in range(L)]


def main():
    sk, pk = o.keygen(IKM)
    hdr, msgs = header(), messages()
    t = time.time()
    sig = o.sign(sk, pk, hdr, msgs)
    print("sign %.1fs" % (time.time() - t), flush=True)
    disclosed = list(range(0, L, 2))
    U = L - len(disclosed)
    t = time.time()
    scalars = o.chacha_random_scalars(SEED, 5 + U)
    proof = o.proof_gen(pk, sig, hdr, PH, msgs, disclosed, scalars)
    print("proof_gen %.1fs" % (time.time() - t), flush=True)
    out = {"L": L, "ikm": IKM.hex(), "sk": sk.hex(), "pk": pk.hex(), "header": hdr.hex(),
           "message_rule": "message i = i as 8-byte big-endian || b'route4-message-fixture!!'",
           "signature": sig.hex(), "proof_seed": SEED.hex(), "presentation_header": PH.hex(),
           "disclosed": "even indexes", "proof_len": len(proof),
           "proof_sha256": hashlib.sha256(proof).hexdigest()}
    with open(os.path.join(HERE, "l65536.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
