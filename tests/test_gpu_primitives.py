"""GPU tests of the reference ABI's primitive entry points (MSM shapes, G1
add, pairing equalities, prepare_b, sign e, ProofGen T2 preparation) through
libcubbs_bbs_hip.so, checked against the CPU oracle and against the
known answers of the reference's native smoke test
(native/cuda/tests/prepare_b_smoke.cpp, fixture tests/golden/native/).

Error convention under test: structural errors fail the call; an undecodable
point or a non-canonical scalar marks its row with code 1.
"""
import hashlib
import json
import os
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KA = json.load(open(os.path.join(HERE, "golden", "native", "prepare_b_known_answer.json")))
INF = b"\xc0" + bytes(47)


def be(x):
    return (x % R_ORDER).to_bytes(32, "big")


def rnd_scalar(rng):
    return be(rng.randrange(R_ORDER))


@pytest.fixture(scope="module")
def pts(oracle):
    rng = random.Random(7)
    g1 = oracle.g1_generator()
    return [oracle.g1_mul(g1, rnd_scalar(rng)) for _ in range(40)]


def test_msm_known_answer_and_route(gpu):
    points = [bytes.fromhex(h) for h in KA["msm_points"]]
    scalars = [bytes.fromhex(h) for h in KA["msm_scalars"]]
    rc, res = gpu.g1_msm([(points, scalars)])
    assert rc == 0
    assert res[0] == (0, 1, bytes.fromhex(KA["expected_b"]))


def test_prepare_b_known_answer(gpu):
    pk = bytes.fromhex(KA["public_key"])
    hdr = bytes.fromhex(KA["header"])
    scs = [bytes.fromhex(h) for h in KA["message_scalars"]]
    rc, res = gpu.prepare_b([(pk, hdr, scs), (pk, hdr, scs)])
    assert rc == 0
    assert res == [(0, bytes.fromhex(KA["expected_b"]))] * 2
    # a 31-byte scalar is a structural error (call level)
    rc, _ = gpu.prepare_b([(pk, hdr, [scs[0], scs[1][:31]] + scs[2:])])
    assert rc == gpu.ERR_INVALID_INPUT
    # header must be the 64-byte bbs-2023 header
    rc, _ = gpu.prepare_b([(pk, hdr[:63], scs)])
    assert rc == gpu.ERR_INVALID_INPUT


def test_msm_items_match_oracle(gpu, oracle, pts):
    rng = random.Random(11)
    items = []
    for n in (0, 1, 2, 3, 5, 17, 40):
        p = [rng.choice(pts) for _ in range(n)]
        s = [rnd_scalar(rng) for _ in range(n)]
        items.append((p, s))
    # identity point, zero scalar, repeated point (doubling inside the sum)
    items.append(([INF, pts[0], pts[1]], [rnd_scalar(rng), be(0), rnd_scalar(rng)]))
    items.append(([pts[2], pts[2]], [be(5), be(R_ORDER - 5)]))
    rc, res = gpu.g1_msm(items)
    assert rc == 0
    for (p, s), (code, route, out) in zip(items, res):
        assert code == 0 and route == 1
        assert out == (oracle.g1_msm(p, s) if p else INF)


def test_msm_row_errors(gpu, pts, oracle):
    bad_point = bytearray(pts[0])
    bad_point[5] ^= 1  # almost surely off the curve
    items = [([pts[0]], [bytes.fromhex(KA["invalid_fr"])]),   # scalar == r
             ([bytes(bad_point)], [be(3)]),
             ([pts[1]], [be(3)])]
    rc, res = gpu.g1_msm(items)
    assert rc == 0
    assert res[0][0] == 1 and res[1][0] == 1
    assert res[2] == (0, 1, oracle.g1_mul(pts[1], be(3)))
    # wrong scalar length: call-level error
    rc, _ = gpu.g1_msm([([pts[0]], [be(3)[:31]])])
    assert rc == gpu.ERR_INVALID_INPUT


def test_large_msm_tree_reduction(gpu, oracle, pts):
    rng = random.Random(3)
    n = 1500
    p = [pts[i % len(pts)] for i in range(n)]
    s = [rnd_scalar(rng) for _ in range(n)]
    rc, res = gpu.g1_msm([(p, s)])
    assert rc == 0 and res[0][0] == 0 and res[0][1] == 3
    assert res[0][2] == oracle.g1_msm(p, s)


@pytest.mark.parametrize("fused,packed", [(False, False), (True, False), (True, True)])
def test_common_points_generators_and_arbitrary(gpu, oracle, pts, fused, packed):
    rng = random.Random(21)
    gens = [oracle.p1()] + oracle.generators(11)   # P1, Q1, H_1..H_10
    groups = []
    for points in (gens, pts[:9]):
        rows = [[rnd_scalar(rng) for _ in points] for _ in range(5)]
        groups.append((points, rows))
    rc, res = gpu.g1_msm_common_points(groups, fused=fused, packed=packed)
    assert rc == 0
    k = 0
    for points, rows in groups:
        for row in rows:
            assert res[k] == (0, 2, oracle.g1_msm(points, row)), k
            k += 1


def test_common_points_fr_boundary_row(gpu, oracle, pts):
    # reference smoke: a scalar equal to r in row 0 rejects row 0 only
    rows = [[bytes.fromhex(KA["invalid_fr"])] + [be(2)] * 3, [be(1), be(2), be(3), be(4)]]
    for fused, packed in ((False, False), (True, True)):
        rc, res = gpu.g1_msm_common_points([(pts[:4], rows)], fused=fused, packed=packed)
        assert rc == 0
        assert res[0][0] == 1
        assert res[1] == (0, 2, oracle.g1_msm(pts[:4], rows[1]))


def test_plus_variable_matches_oracle(gpu, oracle, pts):
    rng = random.Random(5)
    gens = [oracle.p1()] + oracle.generators(6)
    groups = []
    for points in (gens, pts[10:16]):
        rows = []
        for _ in range(4):
            rows.append(([rnd_scalar(rng) for _ in points], rng.choice(pts), rnd_scalar(rng)))
        groups.append((points, rows))
    rc, res = gpu.g1_msm_plus_variable(groups)
    assert rc == 0
    k = 0
    for points, rows in groups:
        for sc, vp, vs in rows:
            assert res[k] == (0, 2, oracle.g1_msm(points + [vp], sc + [vs]))
            k += 1


def test_proof_verify_msm_pair_matches_oracle(gpu, oracle, pts):
    rng = random.Random(9)
    gens = [oracle.p1()] + oracle.generators(8)
    rows = []
    for _ in range(6):
        cs = [rnd_scalar(rng) for _ in gens]
        pp = [rng.choice(pts) for _ in range(3)]
        ps = [rnd_scalar(rng) for _ in range(4)]
        rows.append((cs, pp, ps))
    rc, res = gpu.proof_verify_msm_pair([(gens, rows)])
    assert rc == 0
    for (cs, pp, ps), (code, route, nbytes, t1, t2) in zip(rows, res):
        assert (code, route, nbytes) == (0, 2, 96)
        assert t1 == oracle.g1_msm(pp, ps[:3])
        assert t2 == oracle.g1_msm(gens + [pp[2]], cs + [ps[3]])


def test_dual_msm_and_lanes(gpu, oracle, pts):
    rng = random.Random(13)
    n = 23
    p = [rng.choice(pts) for _ in range(n)]
    a = [rnd_scalar(rng) for _ in range(n)]
    b = [rnd_scalar(rng) for _ in range(n)]
    rc, res = gpu.g1_dual_msm([(p, a, b)])
    assert rc == 0
    assert res[0] == (0, 6, 96, oracle.g1_msm(p, a), oracle.g1_msm(p, b))
    rc, res = gpu.g1_dual_msm([(p, a, b)], lanes=[4])
    assert rc == 0
    code, value, nbytes, L, R = res[0]
    assert (code, value, nbytes) == (0, 6, 4 * 96)
    for j in range(4):
        lo, hi = j * n // 4, (j + 1) * n // 4
        assert L[j] == oracle.g1_msm(p[lo:hi], a[lo:hi])
        assert R[j] == oracle.g1_msm(p[lo:hi], b[lo:hi])
    # lane_count above term_count is a structural error
    rc, _ = gpu.g1_dual_msm([(p[:2], a[:2], b[:2])], lanes=[3])
    assert rc == gpu.ERR_INVALID_INPUT


def test_g1_add(gpu, oracle, pts):
    known = bytes.fromhex(KA["expected_b"])
    p1 = bytes.fromhex(KA["msm_points"][0])
    cases = [(known, p1), (pts[0], pts[0]), (pts[1], INF), (INF, INF)]
    n_neg = oracle.g1_mul(pts[2], be(R_ORDER - 1))
    cases.append((pts[2], n_neg))
    rc, res = gpu.g1_add(cases + [(pts[0][:47], pts[1])])
    assert rc == 0
    for (a, b), (code, out) in zip(cases, res):
        assert code == 0
        assert out == oracle.g1_msm([a, b], [be(1), be(1)]) if a != INF or b != INF else out == INF
    assert res[1][1] == oracle.g1_mul(pts[0], be(2))
    assert res[4][1] == INF
    assert res[-1][0] == 1  # wrong length: per-item code


def test_pairing_eq(gpu, oracle):
    rng = random.Random(17)
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    items, expect = [], []
    for _ in range(7):
        a, b = rng.randrange(1, R_ORDER), rng.randrange(1, R_ORDER)
        P, Q = oracle.g1_mul(g1, be(a)), oracle.g2_mul(g2, be(b))
        items.append((P, Q, oracle.g1_mul(g1, be(a * b)), g2))
        expect.append(True)
        items.append((P, Q, oracle.g1_mul(g1, be(a * b + 1)), g2))
        expect.append(False)
    items.append((INF, g2, g1, b"\xc0" + bytes(95)))   # e(O, G2) == e(G1, O) == 1
    expect.append(True)
    rc, res = gpu.pairing_eq(items)
    assert rc == 0
    assert [c for c, _ in res] == [0] * len(items)
    assert [v for _, v in res] == expect
    for (a1, a2, b1, b2), v in zip(items[:4], expect[:4]):
        assert oracle.pairing_eq(a1, a2, b1, b2) == v
    bad_g2 = bytearray(g2)
    bad_g2[10] ^= 4
    rc, res = gpu.pairing_eq([(g1, bytes(bad_g2), g1, g2)])
    assert rc == 0 and res[0][0] == 1


def test_pairing_eq_aggregate(gpu, oracle):
    rng = random.Random(19)
    g1, g2 = oracle.g1_generator(), oracle.g2_generator()
    # common G2 on each side: e(x_i G1, G2) == e((x_i / s) G1, s G2)
    s = rng.randrange(1, R_ORDER)
    inv_s = pow(s, R_ORDER - 2, R_ORDER)
    QL = oracle.g2_mul(g2, be(s))
    items = []
    for _ in range(20):
        L = oracle.g1_mul(g1, be(rng.randrange(1, R_ORDER)))
        items.append((L, g2, oracle.g1_mul(L, be(inv_s)), QL, rnd_scalar(rng)))
    rc, code, ok = gpu.pairing_eq_aggregate(items)
    assert (rc, code, ok) == (0, 0, True)
    bad = list(items)
    L, Q1, R, Q2, c = bad[7]
    bad[7] = (L, Q1, oracle.g1_mul(R, be(2)), Q2, c)
    assert gpu.pairing_eq_aggregate(bad) == (0, 0, False)
    # varying G2: checked item by item
    mixed = list(items)
    t = rng.randrange(1, R_ORDER)
    mixed[3] = (oracle.g1_mul(g1, be(t)), g2, g1, oracle.g2_mul(g2, be(t)), rnd_scalar(rng))
    assert gpu.pairing_eq_aggregate(mixed) == (0, 0, True)
    assert gpu.pairing_eq_aggregate([]) == (0, 0, True)


def test_proof_gen_t2_prepare(gpu, oracle, pts):
    rng = random.Random(23)
    items = []
    for _ in range(4):
        B, A = rng.choice(pts), rng.choice(pts)
        scs = [rnd_scalar(rng) for _ in range(7)]
        items.append((B, A, scs))
    # multipliers are reduced mod r: a value above r gives the same product
    B, A, scs = items[0]
    big = (int.from_bytes(scs[0], "big") + R_ORDER).to_bytes(32, "big")
    items.append((B, A, [big] + scs[1:]))
    rc, res = gpu.proof_gen_t2_prepare(items)
    assert rc == 0
    for (B, A, s), (code, route, outs) in zip(items, res):
        assert (code, route) == (0, 1)
        d, abar, bbar, t1 = outs
        s = [be(int.from_bytes(x, "big")) for x in s]
        assert d == oracle.g1_mul(B, s[0])
        assert abar == oracle.g1_mul(A, s[1])
        assert bbar == oracle.g1_msm([B, A], [s[2], be(R_ORDER - int.from_bytes(s[3], "big"))])
        assert t1 == oracle.g1_msm([A, B], [s[4], s[5]])
    assert res[4][2] == res[0][2]


def test_sign_e_from_resident_scalars(gpu, oracle):
    rng = random.Random(29)
    msgs_a = [rng.randbytes(rng.randrange(0, 90)) for _ in range(3)]
    msgs_b = [rng.randbytes(40) for _ in range(5)]
    sk, dom_a, dom_b = rnd_scalar(rng), rnd_scalar(rng), rnd_scalar(rng)
    mdst, edst = oracle.MAP_DST, oracle.H2S_DST
    rc, res = gpu.sign_e_from_resident([(sk, dom_a, msgs_a, 0, 3), (sk, dom_b, msgs_b, 3, 5)], mdst, edst)
    assert rc == 0
    for (m, dom), (code, e, ms) in zip(((msgs_a, dom_a), (msgs_b, dom_b)), res):
        sc = [oracle.hash_to_scalar(x, mdst) for x in m]
        assert code == 0 and ms == b"".join(sc)
        assert e == oracle.hash_to_scalar(sk + b"".join(sc) + dom, edst)
    # resident form: scalars of this thread's last messages_to_scalar call
    rc, scal = gpu.messages_to_scalars(msgs_b)
    assert rc == 0
    rc, res = gpu.sign_e_from_resident([(sk, dom_a, None, 1, 3)], mdst, edst)
    assert rc == 0
    assert res[0][1] == oracle.hash_to_scalar(sk + b"".join(scal[1:4]) + dom_a, edst)
    # offset beyond the resident scalars
    rc, _ = gpu.sign_e_from_resident([(sk, dom_a, None, 4, 3)], mdst, edst)
    assert rc == gpu.ERR_INVALID_INPUT


def test_prepare_b_matches_oracle_compute_b(gpu, oracle):
    rng = random.Random(31)
    rc, [(sk, pk)] = gpu.keygen([rng.randbytes(32)])
    hdr = rng.randbytes(64)
    msgs = [rng.randbytes(50) for _ in range(9)]
    scs = [oracle.hash_to_scalar(m, oracle.MAP_DST) for m in msgs]
    rc, res = gpu.prepare_b([(pk, hdr, scs)])
    assert rc == 0 and res[0] == (0, oracle.compute_b(pk, hdr, msgs))
