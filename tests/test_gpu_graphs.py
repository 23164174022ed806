"""Staged batches replayed from a captured HIP graph: every replay's results
and output points must equal the eager (kernel-by-kernel) launch, with eager
work and synchronisations of other batches between the replays, and with two
replays queued back to back.
"""
import hashlib
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def _verify_and_proof_batches(g):
    rng = random.Random(hashlib.sha256(b"graphs").digest())
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(3)])
    assert rc == 0
    sreqs = []
    for i in range(40):
        sk, pk = kps[i % 3]
        L = 1 + (i * 7) % 23
        sreqs.append((sk, pk, rng.randbytes(16), [rng.randbytes(rng.randint(8, 90)) for _ in range(L)]))
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0 and all(c == 0 for c, _ in sigs)
    vreqs = []
    for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs, sigs)):
        if i % 9 == 4:
            m = [b"tampered"] + m[1:]
        vreqs.append((pk, sig, h, m))
    greqs = []
    for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs[:12], sigs[:12])):
        idx = sorted(rng.sample(range(len(m)), len(m) // 3))
        greqs.append((pk, sig, h, rng.randbytes(20), m, idx))
    rc, proofs = g.proof_gen_batch(greqs, seeds=[rng.randbytes(32) for _ in greqs])
    assert rc == 0
    pvreqs = [(pk, proof, h, ph, [m[i] for i in idx], idx, len(m))
              for (pk, sig, h, ph, m, idx), (_, proof) in zip(greqs, proofs)]
    return vreqs, pvreqs


def _msm_items(g, oracle):
    rng = random.Random(17)
    g1 = oracle.g1_generator()
    rc, pts = g.g1_mul([g1] * 1100, [rng.randrange(1, R_ORDER).to_bytes(32, "big") for _ in range(1100)])
    assert rc == 0
    sc = lambda n: [rng.randrange(R_ORDER).to_bytes(32, "big") for _ in range(n)]
    return [(pts, sc(1100)), (pts[:7], sc(7))]


def _run(b, outputs=False):
    b.launch()
    b.sync()
    r = b.results()
    return (r, b.outputs()) if outputs else r


def test_graph_replays_match_eager(gpu, oracle):
    g = gpu
    vreqs, pvreqs = _verify_and_proof_batches(g)
    vb = g.StagedBatch(g.BATCH_VERIFY, vreqs)
    pb = g.StagedBatch(g.BATCH_PROOF_VERIFY, pvreqs)
    mb = g.StagedBatch(g.BATCH_G1_MSM, _msm_items(g, oracle))
    try:
        ev, ep, em = _run(vb), _run(pb), _run(mb, outputs=True)
        assert [v for _, v in ev].count(False) == len([i for i in range(40) if i % 9 == 4])
        assert all(c == 0 and v for c, v in ep)
        for b in (vb, pb, mb):
            b.use_graph(True)
        for rep in range(3):
            assert _run(vb) == ev, rep
            # eager work of another kind between the replays
            assert g.verify_batch(vreqs[:5])[1] == ev[:5]
            assert _run(pb) == ep, rep
            assert _run(mb, outputs=True) == em, rep
        # two replays queued back to back, then one sync
        vb.launch()
        vb.launch()
        vb.sync()
        assert vb.results() == ev
        for b in (vb, pb, mb):
            b.use_graph(False)
        assert _run(vb) == ev and _run(pb) == ep and _run(mb, outputs=True) == em
    finally:
        vb.free()
        pb.free()
        mb.free()


@pytest.mark.parametrize("single_stream,aggregate,nsets", [(False, False, 4), (True, False, 8), (True, True, 8)])
def test_batch_sets_in_flight_match_alone(gpu, oracle, single_stream, aggregate, nsets):
    """Several batches in flight, as in the benchmark: staged Verify +
    ProofVerify sets (forked, or on one stream each) replayed from graphs and
    all launched back to back before any is waited for, over several rounds;
    every set's results equal those of one set run alone and the oracle's
    verdicts."""
    g = gpu
    vreqs, pvreqs = _verify_and_proof_batches(g)
    expect_v = [(0, bool(oracle.verify(*r))) for r in vreqs]
    expect_pv = [(0, bool(oracle.proof_verify(*r))) for r in pvreqs]
    sets = [(g.StagedBatch(g.BATCH_VERIFY, vreqs, single_stream=single_stream, aggregate=aggregate),
             g.StagedBatch(g.BATCH_PROOF_VERIFY, pvreqs, single_stream=single_stream, aggregate=aggregate))
            for _ in range(nsets)]
    try:
        alone = _run(sets[0][0]), _run(sets[0][1])
        assert alone == (expect_v, expect_pv)
        for v, p in sets:
            v.use_graph(True)
            p.use_graph(True)
        for rnd in range(3):
            for v, p in sets:
                v.launch()
                p.launch()
            for v, p in sets:
                v.sync()
                p.sync()
                assert (v.results(), p.results()) == alone, rnd
    finally:
        for v, p in sets:
            v.free()
            p.free()
