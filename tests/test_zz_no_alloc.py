"""The no-allocation variant of messages_to_scalar
(cubbs_hip_messages_to_scalar_noalloc_batch) against the allocating entry
(cubbs_bbs_gpu_messages_to_scalar_batch_v2), through the C ABI: two calls in
a row, the second smaller (the static device array keeps the first call's
contents), equal results and scalars; the MAP DST of the ciphersuite and a
caller DST; the CFRG MapMessageToScalarAsHash vectors; and the call-level
errors (mixed DSTs, an input above the static array's size).
"""
import random

import pytest

from conftest import load_json

pytestmark = pytest.mark.gpu


def _both(g, msgs, dst):
    rc_a, sc_a = g.messages_to_scalars(msgs, dst)
    rc_b, sc_b, res = g.messages_to_scalars_noalloc(msgs, dst)
    assert rc_a == 0 and rc_b == 0
    assert res == [(0, 32)] * len(msgs)
    return sc_a, sc_b


def test_noalloc_matches_allocating_entry_twice(gpu):
    g = gpu
    rng = random.Random(808)
    big = [rng.randbytes(rng.randint(0, 300)) for _ in range(5000)]
    small = [rng.randbytes(rng.randint(0, 90)) for _ in range(37)]
    for msgs in (big, small):
        a, b = _both(g, msgs, g.MAP_DST)
        assert a == b
    dst = b"caller-dst-" + bytes(range(40))
    a, b = _both(g, small[:5], dst)
    assert a == b


def test_noalloc_cfrg_vectors(gpu):
    g = gpu
    fx = load_json("MapMessageToScalarAsHash.json")
    dst = bytes.fromhex(fx["dst"])
    msgs = [bytes.fromhex(c["message"]) for c in fx["cases"]]
    rc, sc, res = g.messages_to_scalars_noalloc(msgs, dst)
    assert rc == 0
    assert [s.hex() for s in sc] == [c["scalar"] for c in fx["cases"]]


def test_noalloc_call_level_errors(gpu):
    import ctypes
    g = gpu
    keep = g._Keep()
    items = (g.MessagesToScalarItem * 2)()
    for i, dst in enumerate((b"dst-a", b"dst-b")):
        items[i].message = keep.slice(b"m")
        items[i].dst = keep.slice(dst)
        _, bb = keep.out(32)
        items[i].scalar_out = bb
    res = (g.Result * 2)()
    assert g.lib().cubbs_hip_messages_to_scalar_noalloc_batch(g.ABI_VERSION, items, ctypes.c_size_t(2), res) == g.ERR_INVALID_INPUT
    # one message larger than the static array: refused before any device work
    huge = bytes(64 << 20)
    items[0].message = keep.slice(huge)
    items[0].dst = keep.slice(b"dst-a")
    assert g.lib().cubbs_hip_messages_to_scalar_noalloc_batch(g.ABI_VERSION, items, ctypes.c_size_t(1), res) == g.ERR_INVALID_INPUT
    assert g.lib().cubbs_hip_messages_to_scalar_noalloc_batch(g.ABI_VERSION + 1, items, ctypes.c_size_t(1), res) == g.ERR_INVALID_INPUT
    assert g.lib().cubbs_hip_messages_to_scalar_noalloc_batch(g.ABI_VERSION, None, ctypes.c_size_t(0), None) == 0
    # the entry still works afterwards
    a, b = _both(g, [b"after", b""], g.MAP_DST)
    assert a == b
