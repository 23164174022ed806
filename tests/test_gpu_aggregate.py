"""Aggregated pairing checks (random linear combination per group of checks
under one key, then per-item checks only for failing groups): every per-item
result must be the same with aggregation off (CUBBS_HIP_AGG_GROUP=0), with
groups of 2 and with groups of 8, and equal the oracle's verdict, for batches
that mix valid items, tampered items spread over groups, a fully tampered
group, undecodable encodings and identity points.
"""
import os
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def _with_group(n, fn):
    old = os.environ.get("CUBBS_HIP_AGG_GROUP")
    os.environ["CUBBS_HIP_AGG_GROUP"] = str(n)
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["CUBBS_HIP_AGG_GROUP"]
        else:
            os.environ["CUBBS_HIP_AGG_GROUP"] = old


def _issuers(g, rng, n):
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(n)])
    assert rc == 0
    return kps


def test_verify_aggregation_matches_individual_checks(gpu, oracle):
    g = gpu
    rng = random.Random(2024)
    kps = _issuers(g, rng, 4)
    sreqs = []
    for i in range(60):
        sk, pk = kps[0] if i < 30 else kps[1 + i % 3]   # one large issuer, three small
        sreqs.append((sk, pk, rng.randbytes(8), [rng.randbytes(rng.randint(1, 60)) for _ in range(1 + i % 7)]))
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    reqs = []
    for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs, sigs)):
        if i in (3, 17, 40) or 8 <= i < 16:          # scattered + one whole group of 8
            m = [b"x"] + m[1:]
        if i == 22:
            sig = sig[:48] + R_ORDER.to_bytes(32, "big")      # undecodable e
        if i == 23:
            sig = b"\xc0" + bytes(47) + sig[48:]               # identity A
        reqs.append((pk, sig, h, m))
    expect = [oracle.verify(*r) for r in reqs]
    runs = {n: _with_group(n, lambda: g.verify_batch(reqs)) for n in (0, 2, 8)}
    base = runs[0]
    assert base[0] == 0
    assert [v for _, v in base[1]] == [bool(e) for e in expect]
    assert base[1][22][0] == 1
    for n in (2, 8):
        assert runs[n] == base, n


def test_proof_verify_aggregation_matches_individual_checks(gpu, oracle):
    g = gpu
    rng = random.Random(77)
    kps = _issuers(g, rng, 3)
    sreqs = []
    for i in range(24):
        sk, pk = kps[i % 3]
        sreqs.append((sk, pk, rng.randbytes(8), [rng.randbytes(30) for _ in range(6)]))
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    greqs = []
    for (sk, pk, h, m), (_, sig) in zip(sreqs, sigs):
        greqs.append((pk, sig, h, rng.randbytes(12), m, sorted(rng.sample(range(6), 2))))
    rc, proofs = g.proof_gen_batch(greqs, seeds=[rng.randbytes(32) for _ in greqs])
    assert rc == 0
    reqs = []
    for i, ((pk, sig, h, ph, m, idx), (_, proof)) in enumerate(zip(greqs, proofs)):
        disc = [m[k] for k in idx]
        if i in (4, 5):
            disc = [b"wrong"] + disc[1:]                      # challenge mismatch
        if i == 9:
            proof = proof[:200] + bytes([proof[200] ^ 1]) + proof[201:]   # a response scalar
        if i == 13:
            proof = b"\xc0" + bytes(47) + proof[48:]          # identity Abar
        reqs.append((pk, proof, h, ph, disc, idx, len(m)))
    runs = {n: _with_group(n, lambda: g.proof_verify_batch(reqs)) for n in (0, 2, 8)}
    base = runs[0]
    assert base[0] == 0
    valid = [v for _, v in base[1]]
    assert valid == [i not in (4, 5, 9, 13) for i in range(24)]
    for n in (2, 8):
        assert runs[n] == base, n


def test_aggregated_staged_batch_graph_replays(gpu, oracle):
    """A staged Verify batch with aggregated groups, replayed from a graph."""
    g = gpu
    rng = random.Random(5)
    kps = _issuers(g, rng, 2)
    sreqs = [(kps[i % 2][0], kps[i % 2][1], b"", [rng.randbytes(20)]) for i in range(40)]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    reqs = [(pk, sig, h, ([b"bad"] if i % 13 == 5 else m)) for i, ((sk, pk, h, m), (_, sig)) in
            enumerate(zip(sreqs, sigs))]
    expect = [(0, i % 13 != 5) for i in range(40)]
    b = _with_group(8, lambda: g.StagedBatch(g.BATCH_VERIFY, reqs))
    try:
        b.launch()
        b.sync()
        assert b.results() == expect
        # 20 items per key in groups of 8, 8, 4; items 5, 18 and 31 are
        # tampered and sit in three different groups: the other three hold
        assert b.group_checks() == (6, 3)
        b.use_graph(True)
        for _ in range(3):
            b.launch()
            b.sync()
            assert b.results() == expect
    finally:
        b.free()


def test_aggregated_groups_hold_for_valid_proofs(gpu, oracle):
    """Every group check of an all-valid ProofVerify batch holds, so no item
    needs its own pairing; one tampered proof fails its group only."""
    g = gpu
    rng = random.Random(31)
    kps = _issuers(g, rng, 2)
    sreqs = [(kps[i % 2][0], kps[i % 2][1], rng.randbytes(8), [rng.randbytes(24) for _ in range(5)])
             for i in range(32)]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    greqs = [(pk, sig, h, rng.randbytes(10), m, [0, 3]) for (sk, pk, h, m), (_, sig) in zip(sreqs, sigs)]
    rc, proofs = g.proof_gen_batch(greqs, seeds=[rng.randbytes(32) for _ in greqs])
    assert rc == 0
    reqs = [(pk, proof, h, ph, [m[0], m[3]], [0, 3], 5)
            for (pk, sig, h, ph, m, idx), (_, proof) in zip(greqs, proofs)]
    b = _with_group(8, lambda: g.StagedBatch(g.BATCH_PROOF_VERIFY, reqs))
    try:
        b.launch()
        b.sync()
        assert b.results() == [(0, True)] * 32
        assert b.group_checks() == (4, 4)
    finally:
        b.free()
    bad = list(reqs)
    pk, proof, h, ph, disc, idx, L = bad[6]
    # Abar of another proof under the same key: a valid point for which the
    # pairing equation fails (and so does the challenge)
    bad[6] = (pk, reqs[8][1][:48] + proof[48:], h, ph, disc, idx, L)
    b = _with_group(8, lambda: g.StagedBatch(g.BATCH_PROOF_VERIFY, bad))
    try:
        b.launch()
        b.sync()
        res = b.results()
        assert [v for _, v in res] == [i != 6 for i in range(32)]
        assert b.group_checks() == (4, 3)
    finally:
        b.free()


def test_aggregate_prepare_flag(gpu, oracle):
    """CUBBS_HIP_BATCH_AGGREGATE at prepare (no environment variable) forms
    groups of 8 per key; results equal the per-item batch's; an unknown flag
    bit is rejected."""
    import ctypes
    g = gpu
    assert os.environ.get("CUBBS_HIP_AGG_GROUP") is None
    rng = random.Random(91)
    kps = _issuers(g, rng, 2)
    sreqs = [(kps[i % 2][0], kps[i % 2][1], b"h", [rng.randbytes(16)]) for i in range(33)]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    reqs = [(pk, sig, h, ([b"bad"] if i == 20 else m)) for i, ((sk, pk, h, m), (_, sig)) in
            enumerate(zip(sreqs, sigs))]
    plain = g.StagedBatch(g.BATCH_VERIFY, reqs)
    agg = g.StagedBatch(g.BATCH_VERIFY, reqs, single_stream=True, aggregate=True)
    try:
        for b in (plain, agg):
            b.launch()
            b.sync()
        assert agg.results() == plain.results() == [(0, i != 20) for i in range(33)]
        # 17 + 16 items per key in groups of at most 8: 8 + 8 (+ one item
        # checked alone) and 8 + 8; item 20 (key 0) fails its group only
        assert plain.group_checks() == (0, 0)
        assert agg.group_checks() == (4, 3)
    finally:
        plain.free()
        agg.free()
    st = ctypes.c_int32(0)
    h = g.lib().cubbs_hip_batch_prepare_flags(g.ABI_VERSION, g.BATCH_VERIFY, None, 0, 4, ctypes.byref(st))
    assert not h and st.value == g.ERR_INVALID_INPUT
