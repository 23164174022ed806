"""Aggregated pairing checks (random linear combination per group of checks
under one key, then per-item checks only for failing groups): every per-item
result must be the same with aggregation off, with groups of 2 and with
groups of 8 (the group size is a prepare flag of the staged batch), and equal
the oracle's verdict, for batches that mix valid items, tampered items spread
over groups, a fully tampered group, undecodable encodings, identity points,
points on the curve outside G1 and non-canonical scalars next to valid items
of the same key.
"""
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def _run(g, kind, reqs, group):
    """(status, per-item results) of one staged batch; group 0 = per-item
    checks, (n, m) = groups of n in supergroups of m groups, (n, m, s) = the
    same with Verify groups' sums formed by product jobs of s signatures."""
    group, sup, straus = (tuple(group) + (0, 0))[:3] if isinstance(group, tuple) else (group, 0, 0)
    b = g.StagedBatch(kind, reqs, aggregate=group if group else False, supergroup=sup, straus=straus)
    try:
        b.launch()
        b.sync()
        return 0, b.results()
    finally:
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
    # (8, 0, s): groups of 8 (9 for the large key's last ones) whose sums
    # are product jobs of s signatures, also with supergroups
    shapes = (2, 8, (8, 4), (2, 3), (8, 0, 1), (8, 0, 2), (8, 0, 3), (8, 0, 4), (8, 4, 4), (2, 3, 2))
    runs = {n: _run(g, g.BATCH_VERIFY, reqs, n) for n in (0,) + shapes}
    base = runs[0]
    assert base[0] == 0
    assert [v for _, v in base[1]] == [bool(e) for e in expect]
    assert base[1][22][0] == 1
    for n in shapes:
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
    runs = {n: _run(g, g.BATCH_PROOF_VERIFY, reqs, n) for n in (0, 2, 8, (2, 4), (4, 2))}
    base = runs[0]
    assert base[0] == 0
    valid = [v for _, v in base[1]]
    assert valid == [i not in (4, 5, 9, 13) for i in range(24)]
    for n in (2, 8, (2, 4), (4, 2)):
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
    b = g.StagedBatch(g.BATCH_VERIFY, reqs, aggregate=8)
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


@pytest.mark.parametrize("straus", [0, 1, 3, 4])
def test_supergroup_levels_gate_groups_and_items(gpu, oracle, straus):
    """Supergroups of up to 4 groups of 8 per key, replayed from a graph:
    the per-item results equal the per-item batch's, a supergroup that holds
    accepts its groups and items without their own checks, and a failing one
    has its groups checked, and the items of failing groups.  straus > 0: the
    groups' sums are product jobs of that many signatures, the items' own
    products run only for the failing group."""
    g = gpu
    rng = random.Random(6)
    kps = _issuers(g, rng, 3)
    # keys 0 and 1: 20 items each (groups of 8, 8, 4: one supergroup of 3
    # groups); key 2: 9 items (one group of 9: a key's last lone item joins
    # the previous group)
    owners = [i % 2 for i in range(40)] + [2] * 9
    sreqs = [(kps[k][0], kps[k][1], b"", [rng.randbytes(20)]) for k in owners]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    bad = {18}                     # key 0: the second item of its second group
    reqs = [(pk, sig, h, ([b"bad"] if i in bad else m)) for i, ((sk, pk, h, m), (_, sig)) in
            enumerate(zip(sreqs, sigs))]
    expect = [(0, i not in bad) for i in range(len(reqs))]
    assert _run(g, g.BATCH_VERIFY, reqs, 0) == (0, expect)
    b = g.StagedBatch(g.BATCH_VERIFY, reqs, aggregate=8, supergroup=4, straus=straus)
    try:
        for graph in (False, True):
            b.use_graph(graph)
            for _ in range(2):
                b.launch()
                b.sync()
                assert b.results() == expect
                # 2 supergroups + 7 groups; key 1's supergroup holds (its 3
                # groups accepted), key 0's fails and two of its groups hold,
                # key 2's lone group holds
                assert b.group_checks() == (9, 7)
                w = b.work()
                assert w["group_checks"] == 9 and w["pair_checks"] == len(reqs)
                # run: 2 supergroup checks, key 0's 3 groups and key 2's
                # group, the 8 items of key 0's failing group
                assert w["pair_checks_run"] == 2 + 4 + 8
                if straus:
                    # groups of 8, 8, 4, 8, 8, 4 and 9 items in jobs of at
                    # most `straus` signatures, and (straus > 1) the failing
                    # group's 8 items' own gated dual products; jobs of one
                    # item are the items' own products
                    jobs = sum(-(-n // straus) for n in (8, 8, 4, 8, 8, 4, 9))
                    gated = 8 if straus > 1 else 0
                    assert w["straus_short_sums"] == jobs + gated
                    assert w["straus_short_terms"] == 49 + gated
                    assert w["glv_products"] == w["agg_products"] == 0
                else:
                    assert w["straus_short_sums"] == 0 and w["glv_products"] == w["agg_products"] == 49
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
    b = g.StagedBatch(g.BATCH_PROOF_VERIFY, reqs, aggregate=8)
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
    b = g.StagedBatch(g.BATCH_PROOF_VERIFY, bad, aggregate=8)
    try:
        b.launch()
        b.sync()
        res = b.results()
        assert [v for _, v in res] == [i != 6 for i in range(32)]
        assert b.group_checks() == (4, 3)
    finally:
        b.free()


def test_aggregate_prepare_flag(gpu, oracle):
    """CUBBS_HIP_BATCH_AGGREGATE at prepare forms groups of 8 per key; results
    equal the per-item batch's; an unknown flag bit, a group size outside
    2..8 and a group size without the aggregate flag are rejected."""
    import ctypes
    g = gpu
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
        # 17 + 16 items per key: 8 + 9 (the last lone item joins the
        # previous group) and 8 + 8; item 20 (key 0) fails its group only
        assert plain.group_checks() == (0, 0)
        assert agg.group_checks() == (4, 3)
    finally:
        plain.free()
        agg.free()
    # an unknown bit, group sizes 1 and 9, a group size without the flag,
    # product jobs of 5 signatures, product jobs without the flag
    for flags in (4, g.BATCH_AGGREGATE | (1 << 8), g.BATCH_AGGREGATE | (9 << 8), 4 << 8,
                  g.BATCH_AGGREGATE | (5 << 24), 2 << 24):
        st = ctypes.c_int32(0)
        h = g.lib().cubbs_hip_batch_prepare_flags(g.ABI_VERSION, g.BATCH_VERIFY, None, 0, flags,
                                                  ctypes.byref(st))
        assert not h and st.value == g.ERR_INVALID_INPUT, flags


def test_aggregated_groups_with_points_outside_g1_and_noncanonical_scalars(gpu, oracle):
    """Inside one key's groups: a Verify A and ProofVerify Abar / Bbar on the
    curve but outside G1 (the group sums apply the GLV endomorphism to them
    while the subgroup check runs beside), a non-canonical e and a
    non-canonical proof response scalar, each next to valid items of the same
    key.  The aggregated results equal the per-item run and the oracle's."""
    g = gpu
    rng = random.Random(612)
    kps = _issuers(g, rng, 2)
    bad_pt = _on_curve_outside_g1(oracle)
    sreqs = [(kps[i % 2][0], kps[i % 2][1], rng.randbytes(8), [rng.randbytes(24) for _ in range(4)])
             for i in range(32)]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    vreqs = []
    for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs, sigs)):
        if i in (2, 9):
            sig = bad_pt + sig[48:]                                   # A outside G1
        if i == 12:
            e = int.from_bytes(sig[48:], "big") + R_ORDER            # e + r, non-canonical
            sig = sig[:48] + e.to_bytes(32, "big")
        vreqs.append((pk, sig, h, m))
    base = _run(g, g.BATCH_VERIFY, vreqs, 0)
    assert [c for c, _ in base[1]] == [1 if i in (2, 9, 12) else 0 for i in range(32)]
    # the oracle's None is the reference's Err: code 1
    assert base[1] == [(0, e) if e is not None else (1, False) for e in (oracle.verify(*r) for r in vreqs)]
    for n in (2, 8, (8, 0, 4), (8, 0, 3), (2, 0, 2)):
        assert _run(g, g.BATCH_VERIFY, vreqs, n) == base, n
    greqs = [(pk, sig, h, rng.randbytes(10), m, [1, 2]) for (sk, pk, h, m), (_, sig) in zip(sreqs, sigs)]
    rc, proofs = g.proof_gen_batch(greqs, seeds=[rng.randbytes(32) for _ in greqs])
    assert rc == 0
    preqs = []
    for i, ((pk, sig, h, ph, m, idx), (_, proof)) in enumerate(zip(greqs, proofs)):
        if i == 3:
            proof = bad_pt + proof[48:]                               # Abar outside G1
        if i == 10:
            proof = proof[:48] + bad_pt + proof[96:]                  # Bbar outside G1
        if i == 17:
            s = int.from_bytes(proof[176:208], "big") + R_ORDER      # r1^ + r, non-canonical
            proof = proof[:176] + s.to_bytes(32, "big") + proof[208:]
        preqs.append((pk, proof, h, ph, [m[k] for k in idx], idx, len(m)))
    base = _run(g, g.BATCH_PROOF_VERIFY, preqs, 0)
    assert [c for c, _ in base[1]] == [1 if i in (3, 10, 17) else 0 for i in range(32)]
    assert base[1] == [(0, e) if e is not None else (1, False)
                       for e in (oracle.proof_verify(*r) for r in preqs)]
    for n in (2, 8):
        assert _run(g, g.BATCH_PROOF_VERIFY, preqs, n) == base, n
