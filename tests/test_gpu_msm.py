"""GPU tests of the Pippenger bucket MSM (item MSMs of at least 1024 terms in
cubbs_bbs_gpu_g1_msm_batch_v2, route id 3), checked against the CPU oracle
at sizes it finishes in seconds and, at 2^16 terms, by linearity: with
P_i = a_i G, sum k_i P_i = (sum a_i k_i mod r) G, one oracle multiplication.

Window widths are forced on the test's thread (cubbs_hip_test_force_shapes)
to exercise the signed digit recoding and the bucket halving at every shape
class.
"""
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu

INF = b"\xc0" + bytes(47)


def be(x):
    return (x % R_ORDER).to_bytes(32, "big")


@pytest.fixture(scope="module")
def multiples(gpu, oracle):
    """4096 points a_i G made on the GPU, a sample checked with the oracle."""
    rng = random.Random(101)
    g1 = oracle.g1_generator()
    a = [rng.randrange(1, R_ORDER) for _ in range(4096)]
    rc, pts = gpu.g1_mul([g1] * len(a), [be(x) for x in a])
    assert rc == 0 and all(p is not None for p in pts)
    for i in rng.sample(range(len(a)), 6):
        assert pts[i] == oracle.g1_mul(g1, be(a[i]))
    return g1, a, pts


def _with_window(g, c, fn):
    """fn() with the Pippenger window width forced on this thread."""
    assert g.force_shapes(pip_c=c) == 0
    try:
        return fn()
    finally:
        g.force_shapes()


@pytest.mark.parametrize("window", [0, 4, 7, 11, 13, 16])
def test_pippenger_matches_oracle(gpu, oracle, multiples, window):
    _, _, pts = multiples
    rng = random.Random(200 + window)
    n = 1024 if window else 2048
    p = pts[:n]
    s = [be(rng.randrange(R_ORDER)) for _ in range(n)]
    # edge terms: identity points, zero / one / r-1 scalars, a repeated point
    # with scalars that put it in the same buckets (doublings inside a bucket)
    p[0], p[1] = INF, INF
    s[2], s[3], s[4] = be(0), be(1), be(R_ORDER - 1)
    p[5] = p[6] = p[7]
    s[5] = s[6] = s[7]
    run = lambda: gpu.g1_msm([(p, s)])
    rc, res = _with_window(gpu, window, run) if window else run()
    assert rc == 0
    assert res[0] == (0, 3, oracle.g1_msm(p, s))


def test_pippenger_several_msms_and_small_items(gpu, oracle, multiples):
    _, _, pts = multiples
    rng = random.Random(5)
    items = []
    for n in (1024, 1500, 3000, 40, 1):
        idx = [rng.randrange(len(pts)) for _ in range(n)]
        items.append(([pts[i] for i in idx], [be(rng.randrange(R_ORDER)) for _ in range(n)]))
    # all terms equal: every entry of a window lands in one bucket
    items.append(([pts[9]] * 1100, [be(123456789)] * 1100))
    # small scalars: only the low windows are non-empty
    items.append((pts[:1200], [be(rng.randrange(1 << 20)) for _ in range(1200)]))
    rc, res = gpu.g1_msm(items)
    assert rc == 0
    for (p, s), (code, route, out) in zip(items, res):
        assert code == 0
        assert route == (3 if len(p) >= 1024 else 1)
        assert out == oracle.g1_msm(p, s)


def test_pippenger_all_zero_and_row_error(gpu, oracle, multiples):
    _, _, pts = multiples
    p = pts[:1024]
    zero = [be(0)] * 1024
    bad = list(zero)
    bad[17] = R_ORDER.to_bytes(32, "big")  # non-canonical scalar: row error
    rc, res = gpu.g1_msm([(p, zero), (p, bad), (p[:3], [be(1)] * 3)])
    assert rc == 0
    assert res[0] == (0, 3, INF)
    assert res[1][0] == 1
    assert res[2] == (0, 1, oracle.g1_msm(p[:3], [be(1)] * 3))


def test_pippenger_2_16_points_linearity(gpu, oracle, multiples):
    g1, a, pts = multiples
    rng = random.Random(77)
    n = 1 << 16
    idx = [rng.randrange(len(pts)) for _ in range(n)]
    k = [rng.randrange(R_ORDER) for _ in range(n)]
    rc, res = gpu.g1_msm([([pts[i] for i in idx], [be(x) for x in k])])
    assert rc == 0 and res[0][0] == 0 and res[0][1] == 3
    total = sum(a[i] * x for i, x in zip(idx, k)) % R_ORDER
    assert res[0][2] == oracle.g1_mul(g1, be(total))


def test_staged_msm_batches_relaunch(gpu, oracle, multiples):
    """Staged MSM kinds: outputs equal the one-shot entries' on every relaunch
    (the Pippenger histogram and cursors are rebuilt per launch)."""
    g1, _, pts = multiples
    rng = random.Random(9)
    items = [(pts[:1300], [be(rng.randrange(R_ORDER)) for _ in range(1300)]),
             (pts[100:105], [be(rng.randrange(R_ORDER)) for _ in range(5)])]
    rc, ref = gpu.g1_msm(items)
    assert rc == 0
    rc, gens = gpu.generators(6)
    assert rc == 0
    groups = [(gens, [[be(rng.randrange(R_ORDER)) for _ in range(6)] for _ in range(3)]),
              (pts[:4], [[be(rng.randrange(R_ORDER)) for _ in range(4)] for _ in range(2)])]
    rc, cref = gpu.g1_msm_common_points(groups, packed=True)
    assert rc == 0
    assert cref[0][2] == oracle.g1_msm(gens, groups[0][1][0])
    sb = gpu.StagedBatch(gpu.BATCH_G1_MSM, items)
    cb = gpu.StagedBatch(gpu.BATCH_G1_MSM_COMMON_POINTS, groups)
    try:
        for _ in range(3):
            sb.launch()
            cb.launch()
            sb.sync()
            cb.sync()
            assert sb.results() == [(0, 3), (0, 1)]
            assert sb.outputs() == [r[2] for r in ref]
            assert cb.results() == [(0, 2)] * 5
            assert cb.outputs() == [r[2] for r in cref]
    finally:
        sb.free()
        cb.free()


def test_autotune_table_and_parity(gpu, oracle, multiples):
    """The measured tuner reports one line per size class, and MSMs planned
    with the tuned shapes still equal the oracle."""
    _, _, pts = multiples
    try:
        rc, rep = gpu.autotune(11)
        assert rc == 0
        lines = rep.strip().split("\n")
        assert [ln.split()[0] for ln in lines] == ["pippenger"] * 2 + ["comb"] * 4
        assert gpu.tune_load(rep) == 0
        rng = random.Random(31)
        items = [(pts[:1024], [be(rng.randrange(R_ORDER)) for _ in range(1024)]),
                 (pts[:2048], [be(rng.randrange(R_ORDER)) for _ in range(2048)])]
        rc, res = gpu.g1_msm(items)
        assert rc == 0
        for (p, s), (code, route, out) in zip(items, res):
            assert (code, route) == (0, 3) and out == oracle.g1_msm(p, s)
        rc, gens = gpu.generators(32)
        rows = [[be(rng.randrange(R_ORDER)) for _ in range(32)] for _ in range(3)]
        rc, cres = gpu.g1_msm_common_points([(gens, rows)], packed=True)
        assert rc == 0 and cres[0][2] == oracle.g1_msm(gens, rows[0])
    finally:
        assert gpu.tune_load("") == 0


def test_variable_base_products_window_edges(gpu, oracle, multiples):
    """One-term MSMs over non-generator points run as single-lane GLV products
    with signed 4-bit windows (k = k1 + k2 lambda, k1 P + k2 phi(P)).  Scalars
    put both halves on the recoding edges (nibbles 7, 8, 9, F; k1 or k2 zero;
    0, small values, r - 1); a batch of 120 items mixes lanes whose top window
    differs."""
    _, _, pts = multiples
    x = 0xD201000000010000
    lam = x * x - 1
    pats = [int(c * 32, 16) for c in "789F"] + [int("80" * 16, 16), int("0F" * 16, 16)]
    scalars = [0, 1, 2, 7, 8, 9, 15, 16, 17, 255, (1 << 64) + 3, R_ORDER - 1, R_ORDER - 2,
               lam, lam - 1, lam + 1, 2 * lam, R_ORDER - lam]
    scalars += [(p + p * lam) % R_ORDER for p in pats] + pats + [(p * lam) % R_ORDER for p in pats]
    rng = random.Random(77)
    scalars += [rng.randrange(R_ORDER) for _ in range(120 - len(scalars))]
    items = [([pts[5 + i]], [be(s)]) for i, s in enumerate(scalars)]
    rc, res = gpu.g1_msm(items)
    assert rc == 0
    for (p, s), (code, _, out) in zip(items, res):
        assert code == 0
        k = int.from_bytes(s[0], "big")
        assert out == (INF if k == 0 else oracle.g1_mul(p[0], s[0])), k
