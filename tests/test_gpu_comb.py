"""GPU tests of the fixed-base comb over the resident generator tables.

Generator slots below 130 (P1, Q1, H_1..H_128) use the wide tables (16
windows of 16-bit signed digits); later slots use the 8-bit tables.  Rows of
140 generators mix both in one comb launch.  The scalars put every window
field on the recoding edges: exactly 2^15 (largest positive digit, last table
entry), 2^15 + 1 (first negative digit), all ones (negative digit 1, then a
carry that turns the next field into a zero digit), r - 1 and 1, and random
values.  Outputs are compared with the oracle's MSM.
"""
import os
import random

import pytest

from conftest import R_ORDER

pytestmark = pytest.mark.gpu


def be(x):
    return (x % R_ORDER).to_bytes(32, "big")


def _fields(field16, top):
    """Scalar with window fields 0..14 equal to field16 and the top field `top`."""
    v = top << 240
    for w in range(15):
        v |= field16 << (16 * w)
    assert v < R_ORDER
    return v


EDGES = [
    _fields(0x8000, 0x1234),
    _fields(0x8001, 0x0001),
    _fields(0xFFFF, 0x0000),
    _fields(0x7FFF, 0x7000),
    _fields(0x0080, 0x0080),   # 8-bit fields at exactly 2^7
    _fields(0x81FF, 0x0000),   # 8-bit fields 0xff then 0x81 with carries
    R_ORDER - 1,
    1,
]


def test_comb_tiers_edge_digits(gpu, oracle):
    n = 140
    rc, gens = gpu.generators(n)
    assert rc == 0
    rng = random.Random(404)
    rows = [[be(e)] * n for e in EDGES]
    rows.append([be(EDGES[i % len(EDGES)]) for i in range(n)])
    rows.append([be(rng.randrange(R_ORDER)) for _ in range(n)])
    rc, res = gpu.g1_msm_common_points([(gens, rows)], packed=True)
    assert rc == 0
    for row, (code, _, out) in zip(rows, res):
        assert code == 0
        assert out == oracle.g1_msm(gens, row)


def test_comb_madds_count_wide_and_narrow_terms(gpu, oracle):
    """A Verify batch with L = 300 messages has 302 fixed-base terms: 130 over
    wide tables (16 additions each) and 172 over 8-bit tables (32 each)."""
    if os.environ.get("CUBBS_HIP_COMB16_GENERATORS"):
        pytest.skip("wide table capacity overridden")
    rng = random.Random(12)
    rc, kps = gpu.keygen([rng.randbytes(32)])
    assert rc == 0
    sk, pk = kps[0]
    msgs = [rng.randbytes(16) for _ in range(300)]
    rc, sigs = gpu.sign_batch([(sk, pk, b"", msgs)])
    assert rc == 0
    b = gpu.StagedBatch(gpu.BATCH_VERIFY, [(pk, sigs[0][1], b"", msgs)])
    try:
        shape = b.shape()
        assert shape["msm_terms"] == 302
        assert shape["comb_madds"] == 130 * 16 + 172 * 32
        b.launch()
        b.sync()
        assert b.results() == [(0, True)]
        assert oracle.verify(pk, sigs[0][1], b"", msgs)
    finally:
        b.free()
