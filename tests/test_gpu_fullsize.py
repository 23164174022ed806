"""Full-size parity of the benchmark's batches against the oracle.

The workload is the benchmark's own (bench.build_workload, fixed seeds):
  * 8192 Verify: signatures from 256 issuer keys, message count log-uniform
    in [1, 128], messages of 32-256 random bytes, 64-byte headers, about 1 %
    tampered (one message replaced);
  * 2048 seeded ProofGen (L = 32 messages of 64 bytes, a random 25 % subset
    disclosed, 64-byte presentation header, blinding scalars from ChaCha20),
    then ProofVerify of all 2048 proofs.
Every proof is compared byte for byte with the oracle's proof for the same
inputs and seed, and every Verify / ProofVerify verdict with the oracle's.
The batches are staged as the benchmark stages them (one stream per batch,
pairing checks aggregated per key in groups of 8, replayed from captured HIP
graphs) with three sets in flight at once, over two rounds of replays.  The
oracle runs on the host cores in threads (its calls release the GIL).
"""
import concurrent.futures as cf
import os
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def workload(gpu):
    sys.path.insert(0, ROOT)
    import bench
    return bench.build_workload(gpu)


@pytest.fixture(scope="module")
def oracle_verdicts(workload, oracle):
    vreqs, _, pvreqs, _, _, _ = workload
    with cf.ThreadPoolExecutor(_threads()) as ex:
        v = list(ex.map(lambda r: oracle.verify(*r), vreqs, chunksize=64))
        p = list(ex.map(lambda r: oracle.proof_verify(*r), pvreqs, chunksize=64))
    return v, p


def test_fullsize_proof_gen_matches_oracle(workload, oracle):
    """All 2048 seeded proofs of the benchmark, byte for byte."""
    _, _, _, _, _, (greqs, seeds, proofs) = workload
    assert len(greqs) == 2048 and len(proofs) == 2048

    def one(i):
        pk, sig, hdr, ph, msgs, idx = greqs[i]
        n = 5 + len(msgs) - len(idx)
        return oracle.proof_gen(pk, sig, hdr, ph, msgs, idx, oracle.chacha_random_scalars(seeds[i], n))

    with cf.ThreadPoolExecutor(_threads()) as ex:
        ref = list(ex.map(one, range(len(greqs)), chunksize=32))
    bad = [i for i in range(len(greqs)) if ref[i] is None or ref[i] != proofs[i]]
    assert not bad, "proofs differing from the oracle: %s" % bad[:10]


def test_fullsize_oracle_agrees_with_ground_truth(workload, oracle_verdicts):
    """The oracle's verdicts are the workload's construction: tampered Verify
    items false, every other item (and every proof) true."""
    _, expect_v, _, _, _, _ = workload
    v, p = oracle_verdicts
    assert v == expect_v
    assert sum(1 for x in expect_v if not x) > 40  # the ~1 % tampered items
    assert p == [True] * len(p)


def test_fullsize_sets_in_flight_match_oracle(gpu, workload, oracle_verdicts):
    g = gpu
    vreqs, _, pvreqs, _, _, _ = workload
    ref_v, ref_p = oracle_verdicts
    want_v = [(0, x) for x in ref_v]
    want_p = [(0, x) for x in ref_p]
    sets = []
    try:
        for _ in range(3):
            sets.append((g.StagedBatch(g.BATCH_VERIFY, vreqs, single_stream=True, aggregate=True),
                         g.StagedBatch(g.BATCH_PROOF_VERIFY, pvreqs, single_stream=True, aggregate=True)))
        for v, p in sets:
            v.use_graph(True)
            p.use_graph(True)
        for rnd in range(2):
            for v, p in sets:
                v.launch()
                p.launch()
            for k, (v, p) in enumerate(sets):
                v.sync()
                p.sync()
                rv, rp = v.results(), p.results()
                bad_v = [i for i in range(len(rv)) if rv[i] != want_v[i]]
                bad_p = [i for i in range(len(rp)) if rp[i] != want_p[i]]
                assert not bad_v, "round %d set %d: Verify items differ: %s" % (rnd, k, bad_v[:10])
                assert not bad_p, "round %d set %d: ProofVerify items differ: %s" % (rnd, k, bad_p[:10])
                groups, passed = v.group_checks()
                # 256 keys, 32 items each on average: groups of 8 hold unless
                # they contain a tampered item
                assert groups >= len(vreqs) // 8 and 0 < passed < groups
                w = v.work()
                assert w["group_checks"] == groups
                assert groups <= w["pair_checks_run"] < groups + w["pair_checks"]
    finally:
        for v, p in sets:
            v.free()
            p.free()
