"""The serving loop under the benchmark's load shape, scaled down: Poisson
arrivals (seeded exponential gaps) of a 40 % Verify / 30 % ProofVerify / 20 %
Sign / 10 % ProofGen mix, 64 issuer keys, message counts log-uniform in
[1, 128] with 60-200-byte messages, a few tampered Verify requests.  Every
request's result and output (signature, seeded proof) must equal the same
request run alone through the batch entry points, and the oracle's verdict or
bytes for it; run with graph-replayed sub-batches, with aggregated checks,
and kernel by kernel.
"""
import concurrent.futures as cf
import math
import os
import random

import pytest

from test_gpu_server import _alone

pytestmark = pytest.mark.gpu


def _L(rng):
    return max(1, min(128, int(math.exp(rng.uniform(0, math.log(129))))))


@pytest.fixture(scope="module")
def load(gpu, oracle):
    g = gpu
    rng = random.Random(20000)
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(64)])
    assert rc == 0

    def msgs():
        return [rng.randbytes(rng.randint(60, 200)) for _ in range(_L(rng))]

    base = [(kps[rng.randrange(64)], rng.randbytes(64), msgs()) for _ in range(48 + 36 + 12)]
    rc, sigs = g.sign_batch([(sk, pk, h, m) for (sk, pk), h, m in base])
    assert rc == 0 and all(c == 0 for c, _ in sigs)
    pool = []
    for i in range(48):
        (sk, pk), h, m = base[i]
        if i % 11 == 5:
            m = [b"tampered"] + m[1:]
        pool.append((g.REQ_VERIFY, (pk, sigs[i][1], h, m), None))
    greqs, seeds = [], []
    for i in range(48, 84):
        (sk, pk), h, m = base[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 4))
        greqs.append((pk, sigs[i][1], h, rng.randbytes(64), m, idx))
        seeds.append(rng.randbytes(32))
    rc, proofs = g.proof_gen_batch(greqs, seeds=seeds)
    assert rc == 0
    for (pk, sig, h, ph, m, idx), (_, proof) in zip(greqs, proofs):
        pool.append((g.REQ_PROOF_VERIFY, (pk, proof, h, ph, [m[i] for i in idx], idx, len(m)), None))
    for _ in range(24):
        sk, pk = kps[rng.randrange(64)]
        pool.append((g.REQ_SIGN, (sk, pk, rng.randbytes(64), msgs()), None))
    for i in range(84, 96):
        (sk, pk), h, m = base[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 4))
        pool.append((g.REQ_PROOF_GEN, (pk, sigs[i][1], h, rng.randbytes(64), m, idx), rng.randbytes(32)))
    starts = {g.REQ_VERIFY: 0, g.REQ_PROOF_VERIFY: 48, g.REQ_SIGN: 84, g.REQ_PROOF_GEN: 108}
    sizes = {g.REQ_VERIFY: 48, g.REQ_PROOF_VERIFY: 36, g.REQ_SIGN: 24, g.REQ_PROOF_GEN: 12}
    kinds = [g.REQ_VERIFY, g.REQ_PROOF_VERIFY, g.REQ_SIGN, g.REQ_PROOF_GEN]
    rate, n = 3000.0, 3000
    t, arrivals, order = 0.0, [], []
    for _ in range(n):
        t += rng.expovariate(rate)
        arrivals.append(int(t * 1e9))
        k = rng.choices(kinds, [0.4, 0.3, 0.2, 0.1])[0]
        order.append(starts[k] + rng.randrange(sizes[k]))
    alone = [_alone(g, *p) for p in pool]

    def ref(p):
        kind, pl, seed = p
        if kind == g.REQ_VERIFY:
            return oracle.verify(*pl)
        if kind == g.REQ_PROOF_VERIFY:
            return oracle.proof_verify(*pl)
        if kind == g.REQ_SIGN:
            return oracle.sign(*pl)
        pk, sig, h, ph, m, idx = pl
        return oracle.proof_gen(pk, sig, h, ph, m, idx, oracle.chacha_random_scalars(seed, 5 + len(m) - len(idx)))

    with cf.ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0))))) as ex:
        refs = list(ex.map(ref, pool))
    return pool, order, arrivals, alone, refs


def test_alone_results_match_oracle(gpu, load):
    g = gpu
    pool, _, _, alone, refs = load
    for (kind, _, _), (rc, code, value, out), r in zip(pool, alone, refs):
        assert rc == 0 and code == 0
        if kind in (g.REQ_VERIFY, g.REQ_PROOF_VERIFY):
            assert bool(value) == r
        else:
            assert out == r


@pytest.mark.parametrize("graphs,aggregate", [(True, False), (True, True), (False, False)])
def test_poisson_load_results_equal_alone(gpu, load, graphs, aggregate):
    g = gpu
    pool, order, arrivals, alone, _ = load
    srv = g.Server(slots=2, graphs=graphs, aggregate=aggregate)
    try:
        rc, res, lat, outs = srv.run_open_loop(pool, order, arrivals, outputs=True)
        assert rc == 0
        bad = [i for i, k in enumerate(order) if (res[i], outs[i]) != (alone[k][1:3], alone[k][3])]
        assert not bad, "requests differing from their alone run: %s" % bad[:10]
        st = srv.stats()
        assert st["fault"] == 0 and st["requests"] == len(order)
        assert st["batches"] < len(order) // 4  # requests were batched
        assert max(lat) > 0
    finally:
        srv.close()
