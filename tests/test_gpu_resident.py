"""Bounded resident state: a child process runs with a key table of 16
entries and an 8-bit comb table budget of 40 generators (set in its
environment before the library loads; the settings are read once per
process), submits batches with more distinct keys and more messages than
those bounds, keeps a staged batch's keys pinned while other batches run, and
checks every result against the oracle while the table sizes, the pinned
entries and the resident bytes stay within their bounds
(tests/resident_child.py)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_resident_state_bounded_and_results_unchanged(gpu):
    env = dict(os.environ, CUBBS_HIP_PK_CACHE_ENTRIES="16",
               CUBBS_HIP_GEN_TABLE_BUDGET_BYTES=str(40 * 32 * 128 * 112),
               CUBBS_HIP_COMB16_GENERATORS="8")
    p = subprocess.run([sys.executable, os.path.join(HERE, "resident_child.py")], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=240)
    assert p.returncode == 0, p.stderr.decode()[-3000:]
    out = json.loads(p.stdout.decode().strip().splitlines()[-1])
    assert out["end"]["pk_entries"] <= 16 and out["end"]["table_slots"] == 40
