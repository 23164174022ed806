"""A Verify batch and a ProofVerify batch staged as one batch (a verify set:
every pipeline stage launches once for the items of both).  Its results must
be, item for item, those of the two batches staged separately and the
oracle's verdicts: per-item checks and aggregated groups (with supergroups),
with tampered, undecodable and host-rejected items in both parts, replayed
from a captured graph, and with either part empty.
"""
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def _results(g, kind, reqs, **kw):
    b = g.StagedBatch(kind, reqs, single_stream=True, **kw)
    try:
        b.launch()
        b.sync()
        return b.results()
    finally:
        b.free()


@pytest.fixture(scope="module")
def mixed(gpu, oracle):
    g = gpu
    rng = random.Random(4242)
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(3)])
    assert rc == 0
    sreqs = [(kps[i % 3][0], kps[i % 3][1], rng.randbytes(16),
              [rng.randbytes(rng.randint(1, 90)) for _ in range(1 + (i * 7) % 23)]) for i in range(45)]
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0
    vreqs = []
    for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs, sigs)):
        if i in (4, 19, 33):
            m = [b"tampered"] + m[1:]
        if i == 27:
            sig = sig[:48] + R_ORDER.to_bytes(32, "big")      # non-canonical e
        if i == 30:
            sig = sig[:40]                                     # rejected on the host
        vreqs.append((pk, sig, h, m))
    # proofs over 12 messages, 3 disclosed
    gs = [(kps[i % 3][0], kps[i % 3][1], rng.randbytes(8), [rng.randbytes(40) for _ in range(12)])
          for i in range(20)]
    rc, psigs = g.sign_batch(gs)
    assert rc == 0
    greqs = [(pk, sig, h, rng.randbytes(10), m, sorted(rng.sample(range(12), 3)))
             for (sk, pk, h, m), (_, sig) in zip(gs, psigs)]
    rc, proofs = g.proof_gen_batch(greqs, seeds=[rng.randbytes(32) for _ in greqs])
    assert rc == 0
    preqs = []
    for i, ((pk, sig, h, ph, m, idx), (_, proof)) in enumerate(zip(greqs, proofs)):
        disc = [m[k] for k in idx]
        if i == 6:
            disc = [b"wrong"] + disc[1:]
        if i == 11:
            proof = proof[:200] + bytes([proof[200] ^ 1]) + proof[201:]
        if i == 14:
            idx = [idx[1], idx[0], idx[2]]                     # not increasing: false on the host
        preqs.append((pk, proof, h, ph, disc, idx, len(m)))
    want_v = [oracle.verify(*r) for r in vreqs]
    want_p = [oracle.proof_verify(*r) for r in preqs]
    return vreqs, preqs, want_v, want_p


@pytest.mark.parametrize("agg", [False, 8, 2, (8, 4)])
def test_verify_set_equals_separate_batches(gpu, mixed, agg):
    g = gpu
    vreqs, preqs, want_v, want_p = mixed
    group, sup = agg if isinstance(agg, tuple) else (agg, 0)
    kw = dict(aggregate=group, supergroup=sup)
    sep = _results(g, g.BATCH_VERIFY, vreqs, **kw) + _results(g, g.BATCH_PROOF_VERIFY, preqs, **kw)
    both = _results(g, g.BATCH_VERIFY_SET, (vreqs, preqs), **kw)
    assert both == sep
    nv = len(vreqs)
    assert both[30][0] == 1 and both[27][0] == 1              # host rejection, undecodable e
    assert both[nv + 14] == (0, False)                         # the reference's Ok(false)
    for i, w in enumerate(want_v):
        if i not in (27, 30):
            assert both[i] == (0, bool(w)), i
    for i, w in enumerate(want_p):
        if i != 14:
            assert both[nv + i] == (0, bool(w)), i


@pytest.mark.parametrize("straus", [0, 4])
def test_verify_set_graph_replays_and_empty_parts(gpu, mixed, straus):
    """Graph replays of a verify set equal the separate batches; with
    straus = 4 the Verify groups' sums are product jobs of 4 signatures (the
    separate Verify batch as well)."""
    g = gpu
    vreqs, preqs, _, _ = mixed
    sep_v = _results(g, g.BATCH_VERIFY, vreqs, aggregate=8, straus=straus)
    sep_p = _results(g, g.BATCH_PROOF_VERIFY, preqs, aggregate=8)
    assert sep_v == _results(g, g.BATCH_VERIFY, vreqs)
    b = g.StagedBatch(g.BATCH_VERIFY_SET, (vreqs, preqs), single_stream=True, aggregate=8, straus=straus)
    try:
        b.use_graph(True)
        for _ in range(3):
            b.launch()
            b.sync()
            assert b.results() == sep_v + sep_p
        groups, passed = b.group_checks()
        assert groups > 0 and 0 < passed < groups
    finally:
        b.free()
    assert _results(g, g.BATCH_VERIFY_SET, (vreqs, []), aggregate=8, straus=straus) == sep_v
    assert _results(g, g.BATCH_VERIFY_SET, ([], preqs), aggregate=8) == sep_p
