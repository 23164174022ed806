"""Serving loop (continuous batching) on the GPU: every result of a mixed,
concurrently submitted request stream must equal the result of the same
request run alone through the batch entry points, and Sign / seeded ProofGen
outputs must be byte-identical.  Also checks that the stream really was
batched (fewer batches than requests) and that request errors stay per item.
"""
import ctypes
import hashlib
import random

import pytest

pytestmark = pytest.mark.gpu


def _pool(g, seed, n_issuers=4, n_verify=24, n_pv=12, n_sign=12, n_pg=8):
    rng = random.Random(hashlib.sha256(seed).digest())
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(n_issuers)])
    assert rc == 0

    def msgs():
        L = max(1, min(40, int(2 ** rng.uniform(0, 5.3))))
        return [rng.randbytes(rng.randint(60, 200)) for _ in range(L)]

    sreqs = [(kps[rng.randrange(n_issuers)], rng.randbytes(16), msgs())
             for _ in range(n_verify + n_pv + n_pg)]
    rc, sigs = g.sign_batch([(sk, pk, h, m) for (sk, pk), h, m in sreqs])
    assert rc == 0 and all(c == 0 for c, _ in sigs)
    pool = []
    for i in range(n_verify):
        (sk, pk), h, m = sreqs[i]
        if i % 7 == 3:  # tampered
            m = list(m)
            m[0] = b"tampered"
        pool.append((g.REQ_VERIFY, (pk, sigs[i][1], h, m), None))
    greqs, gseeds = [], []
    for i in range(n_verify, n_verify + n_pv):
        (sk, pk), h, m = sreqs[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 4))
        greqs.append((pk, sigs[i][1], h, rng.randbytes(64), m, idx))
        gseeds.append(rng.randbytes(32))
    rc, proofs = g.proof_gen_batch(greqs, seeds=gseeds)
    assert rc == 0 and all(c == 0 for c, _ in proofs)
    for (pk, sig, h, ph, m, idx), (_, proof) in zip(greqs, proofs):
        pool.append((g.REQ_PROOF_VERIFY, (pk, proof, h, ph, [m[i] for i in idx], idx, len(m)), None))
    for _ in range(n_sign):
        (sk, pk), h, m = kps[rng.randrange(n_issuers)], rng.randbytes(16), msgs()
        pool.append((g.REQ_SIGN, (sk, pk, h, m), None))
    for i in range(n_verify + n_pv, n_verify + n_pv + n_pg):
        (sk, pk), h, m = sreqs[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 2))
        pool.append((g.REQ_PROOF_GEN, (pk, sigs[i][1], h, rng.randbytes(20), m, idx), rng.randbytes(32)))
    rng.shuffle(pool)
    return pool


def _alone(g, kind, payload, seed):
    if kind == g.REQ_VERIFY:
        rc, r = g.verify_batch([payload])
        return rc, r[0][0], int(r[0][1]), b""
    if kind == g.REQ_PROOF_VERIFY:
        rc, r = g.proof_verify_batch([payload])
        return rc, r[0][0], int(r[0][1]), b""
    if kind == g.REQ_SIGN:
        rc, r = g.sign_batch([payload])
        return rc, r[0][0], 0, r[0][1] or b""
    rc, r = g.proof_gen_batch([payload], seeds=[seed])
    return rc, r[0][0], 0, r[0][1] or b""


def test_server_mixed_results_equal_alone(gpu):
    g = gpu
    pool = _pool(g, b"server-mix")
    alone = [_alone(g, *p) for p in pool]
    assert all(a[0] == 0 and a[1] == 0 for a in alone)
    srv = g.Server(slots=3, max_batch=16)
    try:
        # two passes over the pool, arrivals 150 us apart: later requests arrive
        # while earlier batches are still on the device
        order = list(range(len(pool))) * 2
        arrivals = [150_000 * i for i in range(len(order))]
        rc, res, lat, outs = srv.run_open_loop(pool, order, arrivals, outputs=True)
        assert rc == 0
        for i, k in enumerate(order):
            _, code, value, out = alone[k]
            assert res[i] == (code, value), (i, pool[k][0])
            assert outs[i] == out, (i, pool[k][0])
            assert lat[i] > 0
        st = srv.stats()
        assert st["fault"] == 0 and st["requests"] == len(order)
        assert st["batches"] < len(order)  # requests were batched together
        # the tampered verifies are false, everything else verifies
        assert any(v == 0 for (c, v), k in zip(res, order) if pool[k][0] == g.REQ_VERIFY)
    finally:
        srv.close()


def test_server_submit_wait_and_item_errors(gpu):
    g = gpu
    pool = _pool(g, b"server-submit", n_verify=4, n_pv=2, n_sign=2, n_pg=2)
    srv = g.Server(slots=2)
    try:
        tickets = [srv.submit(k, p, s) for k, p, s in pool]
        bad_pk = (b"\x00" * 95, b"\x00" * 80, b"", [b"m"])
        t_bad = srv.submit(g.REQ_VERIFY, bad_pk)
        for (k, p, s), t in zip(pool, tickets):
            code, value, out, lat = srv.wait(t)
            _, c2, v2, o2 = _alone(g, k, p, s)
            assert (code, value, out) == (c2, v2, o2)
        code, value, out, _ = srv.wait(t_bad)
        assert (code, value) == (g.ERR_INVALID_INPUT, 0)
        # an unknown kind or a missing item is rejected at submission
        for req in (g.Request(7, None, None), g.Request(g.REQ_VERIFY, None, None)):
            assert g.lib().cubbs_hip_server_submit(srv.h, ctypes.byref(req)) == -g.ERR_INVALID_INPUT
    finally:
        srv.close()


def test_server_hbm_budget_bounds_batches(gpu):
    """Admission follows the HBM budget (resident state included): with a
    share per slot below one request's estimate a batch holds one such request
    (a request above the share still runs, alone); with room they batch.  Results equal the
    requests run alone either way."""
    g = gpu
    pool = _pool(g, b"server-budget", n_verify=6, n_pv=4, n_sign=3, n_pg=3)
    alone = [_alone(g, *p) for p in pool]
    order = list(range(len(pool)))
    arrivals = [0] * len(order)  # all pending at once
    # 4 slots x 60 KiB: every Verify / ProofVerify / ProofGen request is above
    # a slot's share (pairing line blocks, product tables) but below the budget
    # the budget covers the library's resident state too
    roomy = g.resident_stats()["resident_bytes"] + 2 * (1 << 30)
    for slots, budget, one_each in ((4, 4 * 60 * 1024, True), (2, roomy, False)):
        srv = g.Server(slots=slots, max_batch=64, hbm_budget=budget)
        try:
            rc, res, lat, outs = srv.run_open_loop(pool, order, arrivals, outputs=True)
            assert rc == 0
            for i, k in enumerate(order):
                _, code, value, out = alone[k]
                assert (res[i], outs[i]) == ((code, value), out), (budget, i)
            st = srv.stats()
            assert st["fault"] == 0 and st["requests"] == len(order)
            if one_each:
                # the 3 Sign requests are estimated small and may share one
                n_sign = sum(1 for k in order if pool[k][0] == g.REQ_SIGN)
                assert len(order) - n_sign + 1 <= st["batches"] <= len(order)
            else:
                assert st["batches"] < len(order)
        finally:
            srv.close()


def test_server_refuses_requests_above_the_budget(gpu):
    """A request whose estimate exceeds the whole HBM budget is refused at
    admission with code 1 (it never reaches the device); the requests around
    it, and later ones, are served normally and the server stays healthy."""
    g = gpu
    pool = _pool(g, b"server-refuse", n_verify=3, n_pv=0, n_sign=3, n_pg=0)
    alone = [_alone(g, *p) for p in pool]
    # 2 x 64 KiB: a Verify request (two pairings' line blocks, ~137 KiB) is
    # above the budget, a Sign request below it
    srv = g.Server(slots=2, max_batch=64, hbm_budget=2 * 64 * 1024)
    try:
        order = list(range(len(pool)))
        rc, res, lat, outs = srv.run_open_loop(pool, order, [0] * len(order), outputs=True)
        for i, k in enumerate(order):
            kind = pool[k][0]
            if kind == g.REQ_VERIFY:
                assert res[i] == (g.ERR_INVALID_INPUT, 0)
            else:
                _, code, value, out = alone[k]
                assert (res[i], outs[i]) == ((code, value), out)
        st = srv.stats()
        assert st["fault"] == 0
        assert st["requests"] == sum(1 for k in order if pool[k][0] != g.REQ_VERIFY)
    finally:
        srv.close()


def test_server_open_loop_rejects_unordered_arrivals(gpu):
    """Arrival times must not decrease; the call is rejected before any
    request is submitted, and the server keeps serving afterwards."""
    g = gpu
    pool = _pool(g, b"server-order", n_verify=2, n_pv=1, n_sign=1, n_pg=1)
    srv = g.Server(slots=2)
    try:
        rc, _, _, _ = srv.run_open_loop(pool, [0, 1, 2], [0, 2_000_000, 1_000_000])
        assert rc == g.ERR_INVALID_INPUT
        assert srv.stats()["requests"] == 0
        alone = [_alone(g, *p) for p in pool]
        rc, res, _, _ = srv.run_open_loop(pool, list(range(len(pool))), [0] * len(pool))
        assert rc == 0 and res == [(c, v) for _, c, v, _ in alone]
    finally:
        srv.close()


def test_server_batcher_window_queue_and_stats(gpu):
    """The request batcher of the reference's W3C service in the loop: a burst
    submitted one by one against a queue of capacity 8 (a submit blocks while
    it is full) and a 3 ms coalescing window for partial batches.  Every
    result equals the request run alone, and the batcher statistics add up:
    the size histogram covers every batch and request, no batch exceeds
    max_batch, nothing stays queued, and the requests waited."""
    g = gpu
    pool = _pool(g, b"server-batcher", n_verify=16, n_pv=8, n_sign=6, n_pg=4)
    alone = [_alone(g, *p) for p in pool]
    srv = g.Server(slots=2, max_batch=6, coalesce_us=3000, queue_capacity=8)
    try:
        tickets = [srv.submit(k, p, s) for k, p, s in pool]
        for t, a in zip(tickets, alone):
            code, value, out, _ = srv.wait(t)
            assert (code, value, out) == a[1:]
        st = srv.stats()
        q = srv.queue_stats(hist_len=8)
        assert st["fault"] == 0 and q["fault"] == 0 and q["queued"] == 0
        assert sum(q["hist"]) == st["batches"]
        assert sum(k * c for k, c in enumerate(q["hist"])) == st["requests"] == len(pool)
        assert q["hist"][0] == 0 and q["hist"][7] == 0 and 1 <= q["max_batch"] <= 6
        assert q["total_wait_ns"] > 0
        assert st["batches"] < len(pool)  # the window gathered requests into batches
    finally:
        srv.close()
