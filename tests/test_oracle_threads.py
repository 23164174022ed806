"""The oracle is called from host threads by the full-size parity tests and
the CPU baseline.  Its generator cache grows while other threads read it:
signatures made concurrently, in a fresh process whose cache starts empty,
with message counts that make the cache grow several times, must equal the
same signatures made one at a time afterwards (CPU)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import concurrent.futures as cf, random, sys
sys.path.insert(0, sys.argv[1])
import oracle_py as o
rng = random.Random(7)
sk, pk = o.keygen(bytes(range(32)))
# message counts up to 600: the cache grows 64 -> 128 -> 256 -> 512 -> 1024
reqs = [(sk, pk, rng.randbytes(8), [rng.randbytes(8) for _ in range(L)])
        for L in sorted(rng.randrange(1, 601) for _ in range(96))]
with cf.ThreadPoolExecutor(16) as ex:
    par = list(ex.map(lambda r: o.sign(*r), reqs))
seq = [o.sign(*r) for r in reqs]
bad = [i for i in range(len(reqs)) if par[i] != seq[i] or par[i] is None]
print("bad", bad)
sys.exit(1 if bad else 0)
"""


def test_oracle_generator_cache_grows_under_concurrent_readers():
    r = subprocess.run([sys.executable, "-c", CHILD, os.path.join(ROOT, "oracle")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
