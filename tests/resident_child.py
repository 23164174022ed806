"""Child process of test_gpu_resident.py: runs with a small key table and a
small 8-bit table budget (set in its environment before the library loads)
and checks results against the oracle while the resident state stays within
its bounds.  Prints one JSON line with the observed statistics."""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "cubbs-cred_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cubbs_hip as g  # noqa: E402
import oracle_py as o  # noqa: E402

PK_CAP = int(os.environ["CUBBS_HIP_PK_CACHE_ENTRIES"])
TABLES = int(os.environ["CUBBS_HIP_GEN_TABLE_BUDGET_BYTES"]) // (32 * 128 * 112)


def main():
    o.lib()
    rng = random.Random(7)
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(40)])
    assert rc == 0
    seen = []

    def snap(tag):
        st = g.resident_stats()
        st["tag"] = tag
        seen.append(st)
        assert st["pk_capacity"] == PK_CAP and st["pk_entries"] <= PK_CAP, st
        assert st["table_slots"] <= TABLES and st["table_max_slots"] == TABLES, st
        return st

    def sign_verify(keys, L, tamper=()):
        sreqs = [(kps[k][0], kps[k][1], rng.randbytes(8), [rng.randbytes(rng.randint(1, 50)) for _ in range(L)])
                 for k in keys]
        rc, sigs = g.sign_batch(sreqs)
        assert rc == 0 and all(c == 0 for c, _ in sigs)
        for r, (_, sig) in zip(sreqs[:3], sigs[:3]):
            assert sig == o.sign(*r), "signature differs from the oracle"
        vreqs = []
        for i, ((sk, pk, h, m), (_, sig)) in enumerate(zip(sreqs, sigs)):
            vreqs.append((pk, sig, h, [b"x"] + m[1:] if i in tamper else m))
        rc, res = g.verify_batch(vreqs)
        assert rc == 0
        assert res == [(0, o.verify(*r)) for r in vreqs]
        assert [v for _, v in res] == [i not in tamper for i in range(len(vreqs))]
        return vreqs

    # 40 distinct keys in one batch: more than the table holds (private copies)
    sign_verify(range(40), 3, tamper=(5, 31))
    base = snap("after 40 keys")
    # rounds over changing key subsets: entries replaced, never more than the cap
    for r in range(6):
        sign_verify([(7 * r + i) % 40 for i in range(10)], 2 + r, tamper=(r,))
        st = snap("round %d" % r)
        assert st["pk_pinned"] == 0 and st["live_batches"] == 0 and st["retired_buffers"] == 0, st
    # a staged batch pins its keys; other batches work around them
    vreqs = sign_verify(range(10), 4)
    staged = g.StagedBatch(g.BATCH_VERIFY, vreqs)
    st = snap("staged pinned")
    assert st["pk_pinned"] == 10 and st["live_batches"] == 1, st
    sign_verify(range(10, 30), 2)          # 20 distinct keys, only 6 free entries: private copies
    sign_verify(range(20, 26), 2)          # 6 keys fit beside the pinned ones
    staged.launch()
    staged.sync()
    assert staged.results() == [(0, True)] * 10
    staged.free()
    st = snap("staged freed")
    assert st["pk_pinned"] == 0 and st["live_batches"] == 0, st
    # more messages than generators with resident tables: private tables
    sign_verify([3, 4], TABLES + 20, tamper=(1,))
    st = snap("long messages")
    assert st["generator_slots"] >= TABLES + 22 and st["table_slots"] == TABLES, st
    # the resident bytes stay bounded: later batches add nothing beyond growth
    for r in range(3):
        sign_verify([(5 * r + i) % 40 for i in range(12)], 5)
    end = snap("end")
    assert end["resident_bytes"] == st["resident_bytes"], (st, end)
    assert end["retired_buffers"] == 0, end
    print(json.dumps({"pk_cap": PK_CAP, "tables": TABLES, "first": base, "end": end}))


if __name__ == "__main__":
    main()
