"""Coalescing of small concurrent primitive calls: 24 host threads each issue
single-item cubbs_bbs_gpu_pairing_eq_batch_v2 and cubbs_bbs_gpu_g1_msm_batch_v2
calls at once (the Rust engine's per-item pattern); every call must get
exactly the result of the same call made alone, valid and invalid items
alike, and the counters must show that calls were combined."""
import concurrent.futures as cf
import ctypes
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def be(x):
    return (x % R_ORDER).to_bytes(32, "big")


def _stats(g):
    v = [ctypes.c_uint64() for _ in range(4)]
    assert g.lib().cubbs_hip_coalesce_stats(*[ctypes.byref(x) for x in v]) == 0
    return [x.value for x in v]


def test_concurrent_single_calls_equal_alone(gpu, oracle):
    g = gpu
    rng = random.Random(4242)
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    pitems = []
    for i in range(48):
        a, b = rng.randrange(1, R_ORDER), rng.randrange(1, R_ORDER)
        P, Q = oracle.g1_mul(g1, be(a)), oracle.g2_mul(g2, be(b))
        pitems.append((P, Q, oracle.g1_mul(g1, be(a * b + (i % 3 == 0))), g2))
    bad = bytearray(g2)
    bad[10] ^= 4
    pitems.append((g1, bytes(bad), g1, g2))          # an undecodable key: code 1 for this call only
    rc, pts = g.g1_mul([g1] * 64, [be(rng.randrange(1, R_ORDER)) for _ in range(64)])
    assert rc == 0
    mitems = [([pts[rng.randrange(64)] for _ in range(n)], [be(rng.randrange(R_ORDER)) for _ in range(n)])
              for n in [1, 2, 5, 9, 17, 33] * 6]
    alone_p = [g.pairing_eq([it]) for it in pitems]
    alone_m = [g.g1_msm([it]) for it in mitems]
    assert all(rc == 0 for rc, _ in alone_p) and all(rc == 0 for rc, _ in alone_m)
    for (rc, res), (a1, a2, b1, b2) in zip(alone_p[:6], pitems[:6]):
        assert res[0] == (0, oracle.pairing_eq(a1, a2, b1, b2))
    for (rc, res), (p, s) in zip(alone_m[:6], mitems[:6]):
        assert res[0][2] == oracle.g1_msm(p, s)
    before = _stats(g)
    jobs = [("p", i) for i in range(len(pitems))] * 3 + [("m", i) for i in range(len(mitems))] * 2
    rng.shuffle(jobs)

    def run(job):
        kind, i = job
        return job, (g.pairing_eq([pitems[i]]) if kind == "p" else g.g1_msm([mitems[i]]))

    with cf.ThreadPoolExecutor(24) as ex:
        out = list(ex.map(run, jobs))
    for (kind, i), got in out:
        assert got == (alone_p[i] if kind == "p" else alone_m[i]), (kind, i)
    after = _stats(g)
    p_comb, p_calls = after[0] - before[0], after[1] - before[1]
    m_comb, m_calls = after[2] - before[2], after[3] - before[3]
    assert p_calls == 3 * len(pitems) and m_calls == 2 * len(mitems)
    assert p_comb < p_calls and m_comb < m_calls  # several calls shared a device call
