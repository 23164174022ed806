"""The maximum-size case of the reference's engine tests
(crates/cubbs-bbs/src/tests.rs:894-1030): one signature over 65,536
messages through all four operations.  Sign must equal the oracle's
signature byte for byte; Verify accepts it and rejects a tampered message;
seeded ProofGen with the 32,768 even indexes disclosed must equal the
oracle's proof (compared by length and SHA-256); ProofVerify accepts it and
rejects a tampered disclosed message.  The expected values come from
tests/golden/max_messages/l65536.json, written by make_vector.py beside it
(the oracle needs minutes at this size, so it does not run here)."""
import hashlib
import json
import os

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

VEC = os.path.join(ROOT, "tests", "golden", "max_messages", "l65536.json")


def test_65536_messages_four_operations(gpu):
    v = json.load(open(VEC))
    L = v["L"]
    sk, pk, hdr = bytes.fromhex(v["sk"]), bytes.fromhex(v["pk"]), bytes.fromhex(v["header"])
    msgs = [i.to_bytes(8, "big") + b"route4-message-fixture!!" for i in range(L)]
    print("L=%d: sign" % L, flush=True)
    rc, sigs = gpu.sign_batch([(sk, pk, hdr, msgs)])
    assert rc == 0
    assert sigs[0] == (0, bytes.fromhex(v["signature"]))
    sig = sigs[0][1]
    print("verify", flush=True)
    rc, res = gpu.verify_batch([(pk, sig, hdr, msgs)])
    assert (rc, res) == (0, [(0, True)])
    bad = list(msgs)
    bad[L // 2] = bad[L // 2] + b"\xff"
    rc, res = gpu.verify_batch([(pk, sig, hdr, bad)])
    assert (rc, res) == (0, [(0, False)])
    print("proof gen", flush=True)
    idx = list(range(0, L, 2))
    ph = bytes.fromhex(v["presentation_header"])
    rc, proofs = gpu.proof_gen_batch([(pk, sig, hdr, ph, msgs, idx)], seeds=[bytes.fromhex(v["proof_seed"])])
    assert rc == 0 and proofs[0][0] == 0
    proof = proofs[0][1]
    assert len(proof) == v["proof_len"] == 272 + 32 * (L - len(idx))
    assert hashlib.sha256(proof).hexdigest() == v["proof_sha256"]
    print("proof verify", flush=True)
    disc = [msgs[i] for i in idx]
    rc, res = gpu.proof_verify_batch([(pk, proof, hdr, ph, disc, idx, L)])
    assert (rc, res) == (0, [(0, True)])
    disc[L // 4] = disc[L // 4] + b"\xff"
    rc, res = gpu.proof_verify_batch([(pk, proof, hdr, ph, disc, idx, L)])
    assert (rc, res) == (0, [(0, False)])
