#!/usr/bin/env python3
"""Benchmark: batched Verify + ProofVerify ops/s on one MI355X.

Workload (one "step" = one pass of the device pipeline over both batches,
inputs resident in HBM):
  * 8192 Verify: signatures from 256 seeded issuer keys, message count
    log-uniform in [1, 128], messages 32..256 random bytes, a 64-byte random
    header per signature, ~1% tampered (a message replaced);
  * 2048 ProofVerify: proofs over L = 32 messages of 64 random bytes, a seeded
    random 25% disclosed, 64-byte presentation header, blinding scalars from
    seeded ChaCha20.
All inputs (keys, signatures, proofs) are produced by the library itself
(keygen, Sign and ProofGen run on the device) from fixed seeds.  The timed
region keeps several staged sets in flight (the pipeline tuning table's
count, 20, one hardware queue each).  A set is a verify set: both batches staged as one device batch,
so every stage launches once for the items of both, on one stream, with the
pairing checks aggregated per key in groups of 8 (random linear
combinations; the Verify groups' B sides as one comb MSM per group over
summed term scalars; per-item checks for the items of failing groups),
replayed from its captured HIP graph.  Each step waits for the set launched
that many steps before, compares its per-item results with the expected
ones, and launches it again; after the last step every set is waited for.
While one set's pairing launches finish, the next sets' decode, hashing and
MSM stages already run.  --separate stages each set as two batches instead.

The roofline object prices the timed step's work with the cost model in
cost_model.json (field multiplications per algorithmic unit, units counted by
the library for each batch: comb additions, GLV products, aggregation
products, decoded points, reduction additions, pairing checks actually
computed) and divides it by the measured step time and the measured
v_mad_u64_u32 peak of the device.  A secondary object gives the same for the
pairing kernel alone on a batch of 8192 per-item checks.

After the timed steps, a serving run pushes an open-loop Poisson stream
(20k req/s, 40% Verify / 30% ProofVerify / 20% Sign / 10% ProofGen, 64 issuers,
L log-uniform in [1, 128], 60-200 byte statements) through the library's
continuous-batching server and reports p50/p99 latency at that fixed load.

Prints one JSON line with the metric, a roofline object for the dominant
kernel (VALU v_mad_u64_u32 bound), the latency object of the serving run, and
the CPU baseline (the oracle's restatement of the reference CPU path, timed on
the host's cores on a bounded sample of the same workload).
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import math
import os
import random
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cubbs-cred_amd"))
# HIP maps streams onto GPU_MAX_HW_QUEUES hardware queues (4 by default), read
# once at initialisation.  The library keeps one work stream per queue.  The
# timed loop keeps its sets in flight on one stream each
# (CUBBS_HIP_BATCH_SINGLE_STREAM), so the measurement runs in a child process
# with BENCH_HW_QUEUES queues, one per set.  Measured with verify sets at 20
# steps: 20 queues and 20 sets 2.70-2.73 ms per step, 16 queues and 12 sets
# 3.07, 18 and 18 3.15; 22 queues and 22 sets ran like 20, but 24 queues
# (20 sets) 4.27.  Sets sharing queues measured slower (24 sets on 20 queues:
# 2.93 ms at 60 steps against 2.49).  The serving loop's slots take 8 streams
# each (2 slots at 16).
BENCH_HW_QUEUES = "20"
SERVING_HW_QUEUES = "16"
# two-batch sets in flight in the timed loop (--separate) when the tuning
# table has no entry: while one set's pairing launches finish, the next
# set's decode / hashing / MSM stages already run
INFLIGHT = 8
# verify sets in flight (one stream each) when the tuning table has no entry
INFLIGHT_SETS = 20

N_VERIFY = 8192
N_ISSUERS = 256
N_PROOFS = 2048
PROOF_L = 32
# algorithmic work model: Fp multiplications per unit of work, and
# v_mad_u64_u32 per Fp multiplication
with open(os.path.join(ROOT, "cost_model.json")) as _f:
    COST = json.load(_f)
FPM_PER_PAIRING_CHECK = COST["fpm"]["pair_check"]
MADS_PER_FPM = COST["mads_per_fpm"]


def work_fpm(w):
    """Field multiplications of one batch launch from its work units."""
    f = COST["fpm"]
    return (w["comb_madds"] * f["comb_madd"] + w["glv_products"] * f["glv_product"] +
            w["agg_products"] * f["agg_product"] + w["decoded_points"] * f["decoded_point"] +
            w["point_adds"] * f["point_add"] + w["pair_checks_run"] * f["pair_check"] +
            w["straus_sums"] * f["straus_chain"] + w["straus_terms"] * f["straus_term"] +
            w["straus_short_sums"] * f["straus_short_chain"] +
            w["straus_short_terms"] * f["straus_short_term"])


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def log_uniform_L(rng):
    return max(1, min(128, int(math.exp(rng.uniform(0, math.log(129))))))


def build_workload(g, seed=b"cubbs-hip/bench"):
    """The benchmark's batches, made by the library from fixed seeds: Verify
    requests and their expected verdicts, ProofVerify requests, the issuer
    keys, whole-op call rates, and the seeded ProofGen requests, seeds and
    proofs the ProofVerify batch was made from."""
    rng = random.Random(hashlib.sha256(seed).digest())
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(N_ISSUERS)])
    assert rc == 0
    # ---- Verify batch
    sreqs = []
    for i in range(N_VERIFY):
        sk, pk = kps[rng.randrange(N_ISSUERS)]
        L = log_uniform_L(rng)
        msgs = [rng.randbytes(rng.randint(32, 256)) for _ in range(L)]
        sreqs.append((sk, pk, rng.randbytes(64), msgs))
    t = time.time()
    rc, sigs = g.sign_batch(sreqs)
    assert rc == 0 and all(c == 0 for c, _ in sigs)
    calls = {"sign_8192_uneven_s": time.time() - t}
    log("signed %d in %.2fs" % (N_VERIFY, time.time() - t))
    vreqs, expect_v = [], []
    for (sk, pk, hdr, msgs), (_, sig) in zip(sreqs, sigs):
        ok = True
        if rng.random() < 0.01:
            msgs = list(msgs)
            msgs[rng.randrange(len(msgs))] = rng.randbytes(32)
            ok = False
        vreqs.append((pk, sig, hdr, msgs))
        expect_v.append(ok)
    # ---- ProofVerify batch
    preqs = []
    for i in range(N_PROOFS):
        sk, pk = kps[rng.randrange(N_ISSUERS)]
        msgs = [rng.randbytes(64) for _ in range(PROOF_L)]
        preqs.append((sk, pk, rng.randbytes(64), msgs))
    rc, psigs = g.sign_batch(preqs)
    assert rc == 0
    greqs, seeds = [], []
    for (sk, pk, hdr, msgs), (_, sig) in zip(preqs, psigs):
        idx = sorted(rng.sample(range(PROOF_L), PROOF_L // 4))
        greqs.append((pk, sig, hdr, rng.randbytes(64), msgs, idx))
        seeds.append(rng.randbytes(32))
    t = time.time()
    rc, proofs = g.proof_gen_batch(greqs, seeds=seeds)
    assert rc == 0 and all(c == 0 for c, _ in proofs)
    calls["proof_gen_2048_s"] = time.time() - t
    log("proofs generated %d in %.2fs" % (N_PROOFS, time.time() - t))
    pvreqs = [(pk, proof, hdr, ph, [msgs[i] for i in idx], idx, len(msgs))
              for (pk, sig, hdr, ph, msgs, idx), (_, proof) in zip(greqs, proofs)]
    # whole-op ABI calls from host buffers (packing, upload and download included)
    abi_ops = {"sign_ops_s": round(N_VERIFY / calls["sign_8192_uneven_s"], 1),
               "proof_gen_ops_s": round(N_PROOFS / calls["proof_gen_2048_s"], 1),
               "note": "one whole-op ABI call each (8192 uneven Sign; 2048 seeded ProofGen at L=32), "
                       "host packing and PCIe included, first call of the process"}
    return vreqs, expect_v, pvreqs, kps, abi_ops, (greqs, seeds, [p for _, p in proofs])


MSM_ROWS = 4096
MSM_TERMS = 32
PIP_POINTS = 1 << 16
R_ORDER = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
G1_GENERATOR = bytes.fromhex(
    "97f1d3a73197d7942695638c4fa9ac0fc3688c4f9774b905a14e3a3f171bac586c55e83ff97a1aeffb3af00adb22c6bb")


def msm_run(g, steps=10, warmup=2, graph=True):
    """Batched G1 MSM kernels alone: 4096 fixed-base MSMs of 32 terms over the
    ciphersuite generators [P1, Q1, H_1..H_30] (comb route, one common-points
    group) and one variable-base MSM of 2^16 points k_i G1 (Pippenger route),
    scalars uniform mod r.  Both staged batches stay resident; a step
    launches both on their own streams and waits for both."""
    rng = random.Random(hashlib.sha256(b"cubbs-hip/msm").digest())
    rc, gens = g.generators(MSM_TERMS)
    assert rc == 0
    rows = [[rng.randrange(R_ORDER).to_bytes(32, "big") for _ in range(MSM_TERMS)]
            for _ in range(MSM_ROWS)]
    rc, pts = g.g1_mul([G1_GENERATOR] * PIP_POINTS,
                       [rng.randrange(1, R_ORDER).to_bytes(32, "big") for _ in range(PIP_POINTS)])
    assert rc == 0 and all(p is not None for p in pts)
    scs = [rng.randrange(R_ORDER).to_bytes(32, "big") for _ in range(PIP_POINTS)]
    cb = g.StagedBatch(g.BATCH_G1_MSM_COMMON_POINTS, [(gens, rows)])
    pb = g.StagedBatch(g.BATCH_G1_MSM, [(pts, scs)])
    del rows, pts, scs
    if graph:
        cb.use_graph(True)
        pb.use_graph(True)
    for _ in range(warmup):
        cb.launch()
        pb.launch()
        cb.sync()
        pb.sync()
    t = time.perf_counter()
    for _ in range(steps):
        cb.launch()
        pb.launch()
        cb.sync()
        pb.sync()
    el = time.perf_counter() - t
    rc_ = cb.results() + pb.results()
    assert all(c == 0 for c, _ in rc_), "MSM rows failed"
    routes = sorted(set(v for _, v in rc_))
    first = cb.outputs() + pb.outputs()
    cb.launch()
    pb.launch()
    cb.sync()
    pb.sync()
    cb.results()
    pb.results()
    assert cb.outputs() + pb.outputs() == first, "MSM outputs differ between launches"
    ct, ck = cb.profile(3)
    pt, pk = pb.profile(3)
    cb.free()
    pb.free()
    return {"workload": "4096 x 32-term fixed-base MSMs over [P1, Q1, H_1..H_30] (comb) + "
                        "one 2^16-point variable-base MSM of k_i G1 (Pippenger), scalars uniform mod r",
            "msm_per_s": round((MSM_ROWS + 1) * steps / el, 1), "ms_per_step": round(el * 1e3 / steps, 3),
            "steps": steps, "routes": routes,
            "comb_4096x32_ms": round(ct, 3), "pippenger_2_16_ms": round(pt, 3),
            "kernel_ms": {"comb": {k: round(v, 3) for k, v in ck.items()},
                          "pippenger": {k: round(v, 3) for k, v in pk.items()}}}


def cpu_baseline(vreqs, pvreqs, budget_s=20.0):
    """The oracle (C restatement of the reference CPU path) on the host cores,
    on a bounded random sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as o
    o.lib()
    threads = max(1, min(16, os.cpu_count() or 1))
    rng = random.Random(1)
    # estimate per-op cost with a few single-thread calls, then size the sample
    t = time.time()
    for r in rng.sample(vreqs, 4):
        o.verify(*r)
    for r in rng.sample(pvreqs, 2):
        o.proof_verify(*r)
    per = (time.time() - t) / 6
    n_total = max(threads * 4, int(budget_s / max(per, 1e-4)))
    n_v = int(n_total * N_VERIFY / (N_VERIFY + N_PROOFS))
    n_p = max(1, n_total - n_v)
    sample = [("v", r) for r in rng.sample(vreqs, min(n_v, len(vreqs)))] + \
             [("p", r) for r in rng.sample(pvreqs, min(n_p, len(pvreqs)))]

    def run(job):
        kind, r = job
        return o.verify(*r) if kind == "v" else o.proof_verify(*r)

    t = time.time()
    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(run, sample))
    el = time.time() - t
    nv = sum(1 for k, _ in sample if k == "v")
    return {"value": round(len(sample) / el, 2), "unit": "ops/s", "cores": threads, "kind": "port",
            "sample": "%d Verify + %d ProofVerify drawn from the benchmark batches (oracle C code, "
                      "%d threads)" % (nv, len(sample) - nv, threads)}


def serving_run(g, rate=20000.0, warm_s=5.0, steady_s=30.0, seed=b"cubbs-hip/serving", slots=0,
                graphs=False, aggregate=False):
    """Open-loop serving load through the library's continuous-batching server:
    Poisson arrivals at `rate` req/s, mix 40% Verify / 30% ProofVerify / 20% Sign
    / 10% ProofGen, 64 issuer keys, message count log-uniform in [1, 128] with
    60-200 byte statements.  Requests are drawn from a pool of distinct
    prepared requests; latency is arrival -> completion, percentiles over the
    steady window (after warm_s)."""
    rng = random.Random(hashlib.sha256(seed).digest())
    rc, kps = g.keygen([rng.randbytes(32) for _ in range(64)])
    assert rc == 0

    def msgs():
        return [rng.randbytes(rng.randint(60, 200)) for _ in range(log_uniform_L(rng))]

    n_pool = {"v": 2048, "pv": 1024, "s": 1024, "pg": 512}
    base = [(kps[rng.randrange(64)], rng.randbytes(64), msgs())
            for _ in range(n_pool["v"] + n_pool["pv"] + n_pool["pg"])]
    rc, sigs = g.sign_batch([(sk, pk, h, m) for (sk, pk), h, m in base])
    assert rc == 0 and all(c == 0 for c, _ in sigs)
    pools = {"v": [], "pv": [], "s": [], "pg": []}
    for i in range(n_pool["v"]):
        (sk, pk), h, m = base[i]
        pools["v"].append((g.REQ_VERIFY, (pk, sigs[i][1], h, m), None))
    off = n_pool["v"]
    greqs, gseeds = [], []
    for i in range(off, off + n_pool["pv"]):
        (sk, pk), h, m = base[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 4))
        greqs.append((pk, sigs[i][1], h, rng.randbytes(64), m, idx))
        gseeds.append(rng.randbytes(32))
    rc, proofs = g.proof_gen_batch(greqs, seeds=gseeds)
    assert rc == 0 and all(c == 0 for c, _ in proofs)
    for (pk, sig, h, ph, m, idx), (_, proof) in zip(greqs, proofs):
        pools["pv"].append((g.REQ_PROOF_VERIFY, (pk, proof, h, ph, [m[i] for i in idx], idx, len(m)), None))
    for _ in range(n_pool["s"]):
        sk, pk = kps[rng.randrange(64)]
        pools["s"].append((g.REQ_SIGN, (sk, pk, rng.randbytes(64), msgs()), None))
    off += n_pool["pv"]
    for i in range(off, off + n_pool["pg"]):
        (sk, pk), h, m = base[i]
        idx = sorted(rng.sample(range(len(m)), len(m) // 4))
        pools["pg"].append((g.REQ_PROOF_GEN, (pk, sigs[i][1], h, rng.randbytes(64), m, idx), rng.randbytes(32)))
    pool = pools["v"] + pools["pv"] + pools["s"] + pools["pg"]
    starts = {"v": 0, "pv": n_pool["v"], "s": n_pool["v"] + n_pool["pv"],
              "pg": n_pool["v"] + n_pool["pv"] + n_pool["s"]}
    kinds = ["v", "pv", "s", "pg"]
    weights = [0.4, 0.3, 0.2, 0.1]
    n = int(rate * (warm_s + steady_s))
    t, arrivals, order = 0.0, [], []
    for _ in range(n):
        t += rng.expovariate(rate)
        arrivals.append(int(t * 1e9))
        k = rng.choices(kinds, weights)[0]
        order.append(starts[k] + rng.randrange(n_pool[k]))
    srv = g.Server(slots=slots, graphs=graphs, aggregate=aggregate)
    try:
        # warm the server (generator tables up to L = 128, issuer keys resident)
        srv.run_open_loop(pool, list(range(0, len(pool), 7)), [0] * len(range(0, len(pool), 7)))
        t0 = time.perf_counter()
        rc, res, lat, _ = srv.run_open_loop(pool, order, arrivals)
        wall = time.perf_counter() - t0
        st = srv.stats()
        qs = srv.queue_stats()
    finally:
        srv.close()
    assert rc == 0, "serving run failed with status %d" % rc
    bad = sum(1 for (c, v), k in zip(res, order)
              if c != 0 or (pool[k][0] in (g.REQ_VERIFY, g.REQ_PROOF_VERIFY) and v != 1))
    assert bad == 0, "%d serving results differ from expectation" % bad
    steady = sorted(l for l, a in zip(lat, arrivals) if a >= warm_s * 1e9)
    pct = lambda q: steady[min(len(steady) - 1, int(q * len(steady)))] / 1e6
    vps = sum(1 for k, a in zip(order, arrivals) if a >= warm_s * 1e9 and pool[k][0] in (0, 1))
    return {"offered_rps": rate, "mix": "40% Verify, 30% ProofVerify, 20% Sign, 10% ProofGen",
            "issuers": 64, "requests": n, "warmup_s": warm_s, "window_s": steady_s,
            "p50_ms": round(pct(0.50), 2), "p99_ms": round(pct(0.99), 2), "max_ms": round(steady[-1] / 1e6, 2),
            "achieved_rps": round(n / wall, 1), "verify_proofverify_rps": round(vps / steady_s, 1),
            "batches": st["batches"], "avg_batch": round(st["requests"] / max(1, st["batches"]), 1),
            "max_batch": qs["max_batch"],
            "avg_queue_wait_ms": round(qs["total_wait_ns"] / max(1, st["requests"]) / 1e6, 3),
            "host_plan_ms_per_batch": st["host_plan_ms_per_batch"], "batch_ms_per_batch": st["batch_ms_per_batch"],
            "host_cpus": len(os.sched_getaffinity(0)),
            "sub_batches": "HIP graph replay (per-slot graph updated per batch)" if graphs else "kernel by kernel",
            "pairing_checks": "per-key groups of 8" if aggregate else "per item",
            "max_concurrent_batches": st["max_concurrent_batches"]}


def serving_child(args):
    """The serving measurement in a child process (its own HIP runtime with
    SERVING_HW_QUEUES hardware queues); returns its latency object."""
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES=SERVING_HW_QUEUES)
    cmd = [sys.executable, os.path.abspath(__file__), "--serving-only",
           "--serving-rate", str(args.serving_rate), "--serving-seconds", str(args.serving_seconds),
           "--serving-slots", str(args.serving_slots)] + \
        (["--serving-graphs"] if args.serving_graphs else []) + \
        (["--serving-aggregate"] if args.serving_aggregate else [])
    p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, check=True)
    lat = json.loads(p.stdout.decode().strip().splitlines()[-1])
    lat["hw_queues"] = int(SERVING_HW_QUEUES)
    return lat


ABI_THREADS = 8
G2_GENERATOR = bytes.fromhex(
    "93e02b6052719f607dacd3a088274f65596bd0d09920b61ab5da61bbdc7f5049334cf11213945d57e5ac7d055d042b7e"
    "024aa2b2f08f0a91260805272dc51051c6e47ad4fa403b02b4510b647ae3d1770bac0326a805bbefd48056c8c121bdb8")


class _PrimitiveVerify:
    """The calls the unchanged Rust engine issues for a Verify batch of mixed
    issuers (crates/cubbs-bbs/src/execution.rs:390-477 and its callers):
    messages_to_scalar over every message; the domain of each item (the
    engine hashes it on the host; here the library's hash_to_scalar export);
    B = P1 + Q1 dom + sum H_j m_j through the common-points MSM, one group per
    message count; the two-term MSMs B * 1 + A * (-e); pairing_eq of
    (A, PK, B - e A, G2).  Every call's items are built once, pointing at the
    previous call's output buffers, so a timed round is the five calls."""

    def __init__(self, g, reqs, gens):
        import ctypes
        self.g, self.c = g, ctypes
        k = self.keep = g._Keep()
        n = len(reqs)
        self.n = n
        msgs = [m for _, _, _, ms in reqs for m in ms]
        self.n_msgs = len(msgs)
        self.scal = ctypes.create_string_buffer(32 * max(1, len(msgs)))
        self.m2s = (g.MessagesToScalarItem * max(1, len(msgs)))()
        base = ctypes.addressof(self.scal)
        for i, m in enumerate(msgs):
            self.m2s[i].message = k.slice(m)
            self.m2s[i].dst = k.slice(g.MAP_DST)
            self.m2s[i].scalar_out = g.ByteBuffer(base + 32 * i, 32, 0)
        self.m2s_res = (g.Result * max(1, len(msgs)))()
        # domain inputs: PK || I2OSP(L, 8) || Q1 || H_1..H_L || api_id || I2OSP(|hdr|, 8) || hdr
        dom = [pk + len(ms).to_bytes(8, "big") + b"".join(gens[1:len(ms) + 2]) + g.API_ID +
               len(h).to_bytes(8, "big") + h for pk, _, h, ms in reqs]
        self.dom_packed, self.dom_offs, self.dom_lens = g._packed(dom)
        self.dom_out = ctypes.create_string_buffer(32 * n)
        dbase = ctypes.addressof(self.dom_out)
        one = k.slice((1).to_bytes(32, "big"))
        # B per item: common-points groups by message count
        self.b_out = ctypes.create_string_buffer(48 * n)
        bbase = ctypes.addressof(self.b_out)
        by_L = {}
        off, m0 = [], 0
        for i, (_, _, _, ms) in enumerate(reqs):
            by_L.setdefault(len(ms), []).append(i)
            off.append(m0)
            m0 += len(ms)
        self.groups = (g.CommonPointsItem * len(by_L))()
        for gi, (L, idx) in enumerate(sorted(by_L.items())):
            self.groups[gi].points = k.slices(gens[:L + 2])
            self.groups[gi].term_count = L + 2
            rows = (g.ByteSlice * ((L + 2) * len(idx)))()
            for r, i in enumerate(idx):
                rows[r * (L + 2)] = one
                rows[r * (L + 2) + 1] = g.ByteSlice(dbase + 32 * i, 32)
                for j in range(L):
                    rows[r * (L + 2) + 2 + j] = g.ByteSlice(base + 32 * (off[i] + j), 32)
            k.refs.append(rows)
            self.groups[gi].scalar_rows = ctypes.cast(rows, ctypes.c_void_p)
            self.groups[gi].output_count = len(idx)
            outs = (g.ByteBuffer * len(idx))(*[g.ByteBuffer(bbase + 48 * i, 48, 0) for i in idx])
            k.refs.append(outs)
            self.groups[gi].point_outs = ctypes.cast(outs, ctypes.c_void_p)
        self.n_groups = len(by_L)
        self.grp_res = (g.Result * n)()
        # B * 1 + A * (-e)
        self.r_out = ctypes.create_string_buffer(48 * n)
        rbase = ctypes.addressof(self.r_out)
        self.msm = (g.G1MsmItem * n)()
        for i, (pk, sig, _, _) in enumerate(reqs):
            e = int.from_bytes(sig[48:80], "big")
            pts = (g.ByteSlice * 2)(g.ByteSlice(bbase + 48 * i, 48), k.slice(sig[:48]))
            scs = (g.ByteSlice * 2)(one, k.slice(((-e) % R_ORDER).to_bytes(32, "big")))
            k.refs += [pts, scs]
            self.msm[i].points = ctypes.cast(pts, ctypes.c_void_p)
            self.msm[i].scalars = ctypes.cast(scs, ctypes.c_void_p)
            self.msm[i].term_count = 2
            self.msm[i].point_out = g.ByteBuffer(rbase + 48 * i, 48, 0)
        self.msm_res = (g.Result * n)()
        g2 = k.slice(G2_GENERATOR)
        self.pair = (g.PairingEqItem * n)()
        for i, (pk, sig, _, _) in enumerate(reqs):
            self.pair[i] = g.PairingEqItem(k.slice(sig[:48]), k.slice(pk), g.ByteSlice(rbase + 48 * i, 48), g2)
        self.pair_res = (g.Result * n)()

    def run(self):
        L, c, A = self.g.lib(), self.c, self.g.ABI_VERSION
        ok = L.cubbs_bbs_gpu_messages_to_scalar_batch_v2(A, self.m2s, c.c_size_t(self.n_msgs), self.m2s_res) == 0
        ok = ok and L.cubbs_bbs_gpu_hash_to_scalar_sha256_batch_for_test_v2(
            A, self.dom_packed, self.dom_offs, self.dom_lens, c.c_size_t(self.n), self.g.H2S_DST,
            c.c_size_t(len(self.g.H2S_DST)), self.dom_out) == 0
        ok = ok and L.cubbs_bbs_gpu_g1_msm_common_points_batch_v2(A, self.groups, c.c_size_t(self.n_groups),
                                                                  self.grp_res) == 0
        ok = ok and L.cubbs_bbs_gpu_g1_msm_batch_v2(A, self.msm, c.c_size_t(self.n), self.msm_res) == 0
        ok = ok and L.cubbs_bbs_gpu_pairing_eq_batch_v2(A, self.pair, c.c_size_t(self.n), self.pair_res) == 0
        if not ok:
            raise SystemExit("primitive sequence call failed")
        return [(self.pair_res[i].code, bool(self.pair_res[i].value)) for i in range(self.n)]


def abi_paths(g, vreqs, expect_v, pvreqs, threads=ABI_THREADS, rounds=3):
    """Throughput of the reference ABI's own entry points fed from host
    buffers (host packing, upload and download included) by `threads` host
    threads, each on its own slice of the benchmark batches: the whole-op
    entries cubbs_bbs_gpu_verify_batch_v2 / proof_verify_batch_v2, and the
    primitive sequence the unchanged Rust engine issues for Verify."""
    import ctypes
    rc, gens = g.generators(130)
    assert rc == 0
    nv, npv = len(vreqs) // threads, len(pvreqs) // threads
    vs = [vreqs[t * nv:(t + 1) * nv] for t in range(threads)]
    ps = [pvreqs[t * npv:(t + 1) * npv] for t in range(threads)]
    keep = g._Keep()
    vitems = [g.build_verify_items(v, keep) for v in vs]
    pitems = [g.build_proof_verify_items(p, keep) for p in ps]
    vres = [(g.Result * nv)() for _ in range(threads)]
    pres = [(g.Result * npv)() for _ in range(threads)]
    L, A = g.lib(), g.ABI_VERSION

    def whole(t):
        a = L.cubbs_bbs_gpu_verify_batch_v2(A, vitems[t], ctypes.c_size_t(nv), vres[t])
        b = L.cubbs_bbs_gpu_proof_verify_batch_v2(A, pitems[t], ctypes.c_size_t(npv), pres[t])
        return a == 0 and b == 0

    with cf.ThreadPoolExecutor(threads) as ex:
        assert all(ex.map(whole, range(threads)))  # warm (tables, keys)
        for t in range(threads):
            got = [bool(vres[t][i].value) for i in range(nv)]
            assert got == expect_v[t * nv:(t + 1) * nv], "whole-op Verify results differ"
            assert all(pres[t][i].code == 0 and pres[t][i].value for i in range(npv))
        t0 = time.perf_counter()
        for _ in range(rounds):
            assert all(ex.map(whole, range(threads)))
        el_whole = time.perf_counter() - t0
        prims = [_PrimitiveVerify(g, v, gens) for v in vs]
        first = list(ex.map(lambda p: p.run(), prims))
        for t in range(threads):
            assert [v for _, v in first[t]] == expect_v[t * nv:(t + 1) * nv], "primitive path results differ"
        t0 = time.perf_counter()
        for _ in range(rounds):
            list(ex.map(lambda p: p.run(), prims))
        el_prim = time.perf_counter() - t0
    return {"threads": threads, "rounds": rounds,
            "whole_op_ops_s": round(threads * (nv + npv) * rounds / el_whole, 1),
            "whole_op": "cubbs_bbs_gpu_verify_batch_v2 (%d uneven items) + proof_verify_batch_v2 (%d) per thread "
                        "from host buffers" % (nv, npv),
            "primitive_verify_ops_s": round(threads * nv * rounds / el_prim, 1),
            "primitive_verify": "per thread, %d Verify as the unchanged Rust engine issues them: messages_to_scalar, "
                                "domain hash_to_scalar, common-points MSM per message count, 2-term MSMs B - e A, "
                                "pairing_eq" % nv}


def per_request_gpu(g, vreqs, n=16):
    """Plain per-request dispatch of the same kernels: batch = 1, synchronous."""
    t = time.time()
    for r in vreqs[:n]:
        rc, _ = g.verify_batch([r])
        assert rc == 0
    return n / (time.time() - t)


def codec_run(n=2048, reps=5):
    """Host throughput of the bbs-2023 proofValue codec (the wire format a
    ProofVerify request arrives in): serialize and parse n derived proofValues
    of the ProofVerify workload's shape (1040-byte proof: L = 32, 24 hidden;
    four label-map entries, 4 + 4 disclosed indexes, 64-byte presentation
    header), timed as direct C-ABI calls with a preallocated arena, one host
    thread."""
    import ctypes
    import proofvalue as pv
    lib = pv.lib()
    rng = random.Random(5)
    vals = [pv.DerivedProofValue(rng.randbytes(1040), {"c14n%d" % i: "b%d" % (3 - i) for i in range(4)},
                                 [0, 1, 2, 3], [9, 14, 20, 27], rng.randbytes(64)) for _ in range(n)]
    strs = [pv.serialize_derived(v).encode() for v in vals]
    # the C structs of the inputs, built once
    keep, cvals = [], []
    for v in vals:
        k = pv._Keep()
        labels = (pv.CLabel * 4)(*[pv.CLabel(k.slice(a.encode()), k.slice(b.encode()))
                                   for a, b in v.label_map.items()])
        man = (ctypes.c_uint64 * 4)(*v.mandatory_indexes)
        sel = (ctypes.c_uint64 * 4)(*v.selective_indexes)
        cvals.append(pv.CDerived(k.slice(v.bbs_proof), labels, 4, man, 4, sel, 4, k.slice(v.presentation_header)))
        keep.append((k, labels, man, sel))
    need = ctypes.c_size_t(0)
    out_buf = ctypes.create_string_buffer(4096)
    arena = ctypes.create_string_buffer(8192)
    cout = pv.CDerived()
    ser, par = lib.cubbs_proofvalue_serialize_derived, lib.cubbs_proofvalue_parse_derived
    best_s = best_p = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        for c in cvals:
            if ser(ctypes.byref(c), out_buf, 4096, ctypes.byref(need)) != 0:
                raise SystemExit("proofValue serialize failed")
        best_s = min(best_s, time.perf_counter() - t)
        t = time.perf_counter()
        for sv in strs:
            if par(sv, len(sv), arena, 8192, ctypes.byref(need), ctypes.byref(cout)) != 0:
                raise SystemExit("proofValue parse failed")
        best_p = min(best_p, time.perf_counter() - t)
    return {"items": n, "proof_value_chars": len(strs[0]), "threads": 1,
            "serialize_derived_per_s": round(n / best_s), "parse_derived_per_s": round(n / best_p),
            "timing": "best of %d passes, direct C-ABI calls (ctypes call overhead included)" % reps}


PROFILE_STEP = os.path.join(ROOT, "profiles", "step_verifyset_verify8192_proofverify2048_aggregated.md")
PROFILE_PAIRING = os.path.join(ROOT, "profiles", "pairing_verify8192_per_item.md")


def profiled(path, key):
    """A figure from a committed rocprofv3 summary (the line `key: value`),
    None when the summary or the line is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None


def make_set(g, vreqs, pvreqs, combined, **kw):
    """One staged set of the step's work: a verify set (the Verify and the
    ProofVerify batch staged as one batch: every stage launches once for
    both), or the two batches staged separately; single stream each."""
    if combined:
        return [g.StagedBatch(g.BATCH_VERIFY_SET, (vreqs, pvreqs), single_stream=True, **kw)]
    return [g.StagedBatch(g.BATCH_VERIFY, vreqs, single_stream=True, **kw),
            g.StagedBatch(g.BATCH_PROOF_VERIFY, pvreqs, single_stream=True, **kw)]


def check_set(batches, expect_v):
    """The set's results: Verify verdicts as constructed, every proof valid."""
    res = [r for b in batches for r in b.results()]
    nv = len(expect_v)
    if [x for _, x in res[:nv]] != expect_v or not all(c == 0 and x for c, x in res[nv:]):
        raise SystemExit("results differ from the expected ones")
    return tuple(b.results_raw() for b in batches)


def run_in_flight(sets, steps, expect):
    """Seconds for `steps` steps with the staged sets in flight round-robin:
    a set is launched again once its previous results are back and checked
    against `expect`."""
    n = len(sets)

    def finish(i):
        for b in sets[i]:
            b.sync()
        if tuple(b.results_raw() for b in sets[i]) != expect[i]:
            raise SystemExit("results of a timed step differ from the expected ones")

    t = time.perf_counter()
    for step in range(steps):
        i = step % n
        if step >= n:
            finish(i)
        for b in sets[i]:
            b.launch()
    for step in range(max(0, steps - n), steps):
        finish(step % n)
    return time.perf_counter() - t


# pipeline tuning table (bench.py --tune-pipeline), loaded by the headline run
PIPELINE_TUNING = os.path.join(ROOT, "cubbs-cred_amd", "pipeline_tuning.json")
TUNE_CLASSES = ((2048, 512), (N_VERIFY, N_PROOFS))
TUNE_GROUPS = (4, 8)
TUNE_SUPER = (0, 4)
TUNE_CHUNKS = (2, 4)
# sets in flight per staging form: verify sets take one stream each, the
# separate form two (BENCH_HW_QUEUES hardware queues)
TUNE_INFLIGHT = {True: (12, 16, 20), False: (8, 10)}
TUNE_STEPS = 20  # the headline's timed steps: pipeline fill and drain included


def tune_pipeline(g, vreqs, expect_v, pvreqs, graph=True):
    """Measures every (staging form, aggregation group, groups per
    supergroup, comb terms per lane, sets in flight) candidate for each size
    class (Verify, ProofVerify items per set, the benchmark's requests cut to
    that size) as the headline runs them: single-stream sets replayed from
    graphs, several in flight, results checked.  Returns the table: all
    candidates, the fastest per class, and the library's pipeline lines for
    cubbs_hip_tune_load."""
    classes, lines = [], []
    for nv, npv in TUNE_CLASSES:
        vr, pr, ev = vreqs[:nv], pvreqs[:npv], expect_v[:nv]
        cands = []
        for comb, grp, sup, ch in [(f, a, b, c) for f in (True, False) for a in TUNE_GROUPS
                                   for b in TUNE_SUPER for c in TUNE_CHUNKS]:
            sets = [make_set(g, vr, pr, comb, aggregate=grp, comb_chunk=ch, supergroup=sup)
                    for _ in range(max(TUNE_INFLIGHT[comb]))]
            if graph:
                for st in sets:
                    for b in st:
                        b.use_graph(True)
            for st in sets:
                for b in st:
                    b.launch()
            expect = []
            for st in sets:
                for b in st:
                    b.sync()
                expect.append(check_set(st, ev))
            for n in TUNE_INFLIGHT[comb]:
                run_in_flight(sets[:n], n, expect)  # warm-up
                ms = min(run_in_flight(sets[:n], TUNE_STEPS, expect) for _ in range(3)) * 1e3 / TUNE_STEPS
                cands.append({"verify_set": comb, "agg_group": grp, "agg_super": sup, "comb_chunk": ch,
                              "sets_in_flight": n, "ms_per_step": round(ms, 4)})
                log("tune %d+%d %s group=%d super=%d chunk=%d inflight=%d: %.3f ms/step"
                    % (nv, npv, "set" if comb else "separate", grp, sup, ch, n, ms))
            for st in sets:
                for b in st:
                    b.free()
        best = min(cands, key=lambda c: c["ms_per_step"])
        classes.append({"verify_items": nv, "proof_verify_items": npv, "best": best, "candidates": cands})
        for kind, items in (("verify", nv), ("proof_verify", npv)):
            lines.append("pipeline kind=%s log2_items=%d chunk=%d group=%d super=%d" %
                         (kind, round(math.log2(items)), best["comb_chunk"], best["agg_group"],
                          best["agg_super"]))
    return {"description": "Measured Verify/ProofVerify pipeline shapes per size class (items per set): "
                           "staging form (verify_set: both batches staged as one batch, else two "
                           "batches), aggregation group size, groups per supergroup check (0: none), "
                           "fixed-base comb terms per lane, staged sets in flight; single-stream sets "
                           "replayed from HIP graphs, results checked",
            "command": "python bench.py --tune-pipeline cubbs-cred_amd/pipeline_tuning.json",
            "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
            "classes": classes, "table": lines}


TUNE_STRAUS = (0, 1)


def tune_straus(g, vreqs, expect_v, pvreqs, table, graph=True):
    """Refines a pipeline table: for each size class, at its best shape, times
    the Verify groups' sums as per-item products (straus 0) and as product
    jobs of one signature (straus 1, one table per point for both sums) and
    keeps the faster.  Returns the updated table."""
    lines = []
    for c in table["classes"]:
        nv, npv, best = c["verify_items"], c["proof_verify_items"], dict(c["best"])
        vr, pr, ev = vreqs[:nv], pvreqs[:npv], expect_v[:nv]
        n = best["sets_in_flight"]
        timed = {}
        for st in TUNE_STRAUS:
            sets = [make_set(g, vr, pr, best["verify_set"], aggregate=best["agg_group"],
                             supergroup=best["agg_super"], comb_chunk=best["comb_chunk"], straus=st)
                    for _ in range(n)]
            for one in sets:
                for b in one:
                    b.use_graph(graph)
                    b.launch()
            expect = []
            for one in sets:
                for b in one:
                    b.sync()
                expect.append(check_set(one, ev))
            run_in_flight(sets, n, expect)  # warm-up
            ms = min(run_in_flight(sets, TUNE_STEPS, expect) for _ in range(5)) * 1e3 / TUNE_STEPS
            timed[st] = round(ms, 4)
            c["candidates"].append(dict(best, agg_straus=st, ms_per_step=timed[st]))
            log("tune %d+%d straus=%d: %.3f ms/step" % (nv, npv, st, ms))
            for one in sets:
                for b in one:
                    b.free()
        st = min(timed, key=timed.get)
        c["best"] = dict(best, agg_straus=st, ms_per_step=timed[st])
        for kind, items in (("verify", nv), ("proof_verify", npv)):
            lines.append("pipeline kind=%s log2_items=%d chunk=%d group=%d super=%d straus=%d" %
                         (kind, round(math.log2(items)), best["comb_chunk"], best["agg_group"],
                          best["agg_super"], st if kind == "verify" else 0))
    table["table"] = lines
    note = ("; then, at each class's best shape, the Verify groups' sums as per-item products "
            "(agg_straus 0) or product jobs of one signature (1)")
    if note not in table["description"]:
        table["description"] += note
    step = " && python bench.py --tune-straus cubbs-cred_amd/pipeline_tuning.json"
    if step not in table["command"]:
        table["command"] += step
    return table


def load_pipeline_tuning(g):
    """Installs the committed pipeline table in the library and returns the
    choice for the headline's size class (None without a table)."""
    if not os.path.exists(PIPELINE_TUNING):
        return None
    with open(PIPELINE_TUNING) as f:
        table = json.load(f)
    rc = g.tune_load("\n".join(table["table"]))
    if rc != 0:
        raise SystemExit("pipeline tuning table rejected (rc=%d)" % rc)
    for c in table["classes"]:
        if (c["verify_items"], c["proof_verify_items"]) == (N_VERIFY, N_PROOFS):
            return dict(c["best"])
    return None


def mark():
    """A one-block probe kernel launched (and waited for) just outside the
    timed region, so that a kernel trace of this run can find the region's
    dispatches between the two marker dispatches."""
    import cubbs_hip as g
    g.fp_mul_bench(64, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop latency run")
    ap.add_argument("--serving-rate", type=float, default=20000.0)
    ap.add_argument("--serving-seconds", type=float, default=30.0)
    ap.add_argument("--serving-slots", type=int, default=0, help="server slots (0 = library default)")
    ap.add_argument("--serving-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--serving-graphs", action="store_true",
                    help="serving sub-batches captured into per-slot HIP graphs and replayed (default: kernel by "
                         "kernel, which measured lower latency)")
    ap.add_argument("--serving-aggregate", action="store_true",
                    help="serving Verify / ProofVerify sub-batches with aggregated pairing checks")
    ap.add_argument("--profile-only", action="store_true", help="a few launches, no baselines")
    ap.add_argument("--no-msm", action="store_true", help="skip the batched-MSM measurement")
    ap.add_argument("--msm-only", action="store_true",
                    help="only the batched-MSM measurement (after the measured tuning): for profiling")
    ap.add_argument("--no-graph", action="store_true", help="launch kernel by kernel instead of HIP graphs")
    ap.add_argument("--no-autotune", action="store_true", help="skip the measured launch-shape tuning")
    ap.add_argument("--no-aggregate", action="store_true",
                    help="per-item pairing checks only (no random-linear-combination group checks)")
    ap.add_argument("--separate", action="store_true",
                    help="stage each set as a Verify and a ProofVerify batch instead of one verify set")
    ap.add_argument("--fork", action="store_true",
                    help="staged batches fork their decode/product branch onto a second stream "
                         "(lower latency per batch, two hardware queues each)")
    ap.add_argument("--agg-group", type=int, default=0, metavar="N",
                    help="aggregation group size of the timed sets (2..8; 0: the tuning table's)")
    ap.add_argument("--agg-super", type=int, default=0, metavar="M",
                    help="groups per supergroup check of the timed sets (2..8; 0: the tuning table's)")
    ap.add_argument("--comb-chunk", type=int, default=0, metavar="N",
                    help="fixed-base comb terms per lane of the timed sets (0: the tuning table's)")
    ap.add_argument("--agg-straus", type=int, default=0, metavar="N",
                    help="signatures per product job of an aggregated Verify group's sums (1..4; 0: the "
                         "tuning table's, else two products per item)")
    ap.add_argument("--inflight", type=int, default=0,
                    help="staged batch sets in flight in the timed loop (1 = launch, wait, repeat; "
                         "default: the pipeline tuning table's choice, else %d)" % INFLIGHT)
    ap.add_argument("--tune-straus", metavar="PATH", default="",
                    help="time per-item products against one-signature product jobs at the committed "
                         "table's best shapes and write the refined table to PATH")
    ap.add_argument("--tune-pipeline", metavar="PATH", default="",
                    help="measure the Verify/ProofVerify pipeline shapes (aggregation group, comb "
                         "terms per lane, sets in flight) per size class and write the table to PATH")
    ap.add_argument("--no-pipeline-tuning", action="store_true",
                    help="ignore the committed pipeline tuning table")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(7)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": n}))
        return
    if args.serving_only:
        import cubbs_hip as g
        lat = serving_run(g, rate=args.serving_rate, steady_s=args.serving_seconds, slots=args.serving_slots,
                          graphs=args.serving_graphs, aggregate=args.serving_aggregate)
        print(json.dumps(lat), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent requests shard by issuer; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < int(BENCH_HW_QUEUES):
        # Nothing touches the GPU in this process.  The measurements run in
        # child processes with enough hardware queues, one after the other:
        # the throughput child exits before the serving child starts, so the
        # serving run never shares the device with another process's queues.
        import subprocess
        env = dict(os.environ, GPU_MAX_HW_QUEUES=BENCH_HW_QUEUES)
        cmd = [sys.executable, os.path.abspath(__file__)] + sys.argv[1:] + ["--no-serving"]
        p = subprocess.run(cmd, env=env, stdout=subprocess.PIPE)
        if p.returncode:
            sys.stdout.write(p.stdout.decode())
            sys.exit(p.returncode)
        out = json.loads(p.stdout.decode().strip().splitlines()[-1])
        if not args.no_serving and not args.profile_only and not args.msm_only and not args.tune_pipeline \
                and not args.tune_straus:
            t = time.time()
            out["latency"] = serving_child(args)
            log("serving run %.1fs: %s" % (time.time() - t, out["latency"]))
            if not args.serving_graphs:
                # the same load with graph-replayed sub-batches, shorter window
                g_args = argparse.Namespace(**dict(vars(args), serving_graphs=True,
                                                   serving_seconds=min(10.0, args.serving_seconds)))
                t = time.time()
                out["latency_graphs"] = serving_child(g_args)
                log("serving run (graphs) %.1fs: %s" % (time.time() - t, out["latency_graphs"]))
        print(json.dumps(out), flush=True)
        return
    import cubbs_hip as g
    rc, kind, hw = g.backend_info()
    if rc != 0 or not hw:
        raise SystemExit("no usable HIP device (backend_info rc=%d)" % rc)
    if args.msm_only:
        out = {"workload": "batched MSMs only"}
        if not args.no_autotune:
            rc, rep = g.autotune(16)
            if rc != 0:
                raise SystemExit("autotune failed rc=%d" % rc)
            out["tuning"] = rep.strip().split("\n")
        out["msm"] = msm_run(g, steps=args.steps, warmup=args.warmup, graph=not args.no_graph)
        print(json.dumps(out), flush=True)
        return
    t0 = time.time()
    vreqs, expect_v, pvreqs, kps, abi_ops, _ = build_workload(g)
    log("workload built in %.1fs" % (time.time() - t0))
    if args.tune_straus:
        with open(PIPELINE_TUNING) as f:
            table = tune_straus(g, vreqs, expect_v, pvreqs, json.load(f), graph=not args.no_graph)
        with open(args.tune_straus, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
        print(json.dumps({"pipeline_tuning": [c["best"] for c in table["classes"]]}), flush=True)
        return
    if args.tune_pipeline:
        table = tune_pipeline(g, vreqs, expect_v, pvreqs, graph=not args.no_graph)
        with open(args.tune_pipeline, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
        print(json.dumps({"pipeline_tuning": table["classes"]}), flush=True)
        return
    tuned = None if args.no_pipeline_tuning else load_pipeline_tuning(g)
    combined = not args.separate and not args.fork and (tuned.get("verify_set", True) if tuned else True)
    nset = max(1, args.inflight or (tuned["sets_in_flight"] if tuned else
                                     (INFLIGHT_SETS if combined else INFLIGHT)))
    single, agg = not args.fork, not args.no_aggregate
    straus = args.agg_straus if agg else 0
    # the product jobs in use: the flag's, else the table's when the group
    # size is left to the table too
    straus_used = (straus or ((tuned or {}).get("agg_straus", 0) if not args.agg_group else 0)) if agg else 0
    # the supergroup level in use: the flag's, else the table's when the
    # group size is left to the table too
    sup_used = args.agg_super or ((tuned or {}).get("agg_super", 0) if not args.agg_group else 0)
    if combined:
        grp = (args.agg_group or True) if agg else False
        sets = [make_set(g, vreqs, pvreqs, True, aggregate=grp, straus=straus, comb_chunk=args.comb_chunk,
                         supergroup=args.agg_super) for _ in range(nset)]
    else:
        sets = [[g.StagedBatch(g.BATCH_VERIFY, vreqs, single_stream=single, aggregate=agg, straus=straus),
                 g.StagedBatch(g.BATCH_PROOF_VERIFY, pvreqs, single_stream=single, aggregate=agg)]
                for _ in range(nset)]
    if not args.no_graph:
        # each batch's whole launch sequence replayed from one captured HIP graph
        for st in sets:
            for b in st:
                b.use_graph(True)
    log("staged %d set(s)%s" % (nset, " (verify sets)" if combined else ""))
    for _ in range(max(1, args.warmup)):
        for st in sets:
            for b in st:
                b.launch()
        for st in sets:
            for b in st:
                b.sync()
    expect = [check_set(st, expect_v) for st in sets]

    mark()
    el = run_in_flight(sets, args.steps, expect)
    mark()
    ms_per_step = el * 1e3 / args.steps
    ops = N_VERIFY + N_PROOFS
    # the algorithmic work of one step: the batches of a set (every set holds
    # the same items; pair_checks_run counts the checks of their last launch)
    step_work = [b.work() for b in sets[0]]
    step_fpm = sum(work_fpm(w) for w in step_work)
    msm_terms = sum(b.shape()["msm_terms"] for b in sets[0])
    value = ops * args.steps / el
    for st in sets[1:]:
        for b in st:
            b.free()
    log("timed %d steps (%d in flight): %.2f ms/step, %.0f ops/s" % (args.steps, nset, ms_per_step, value))
    # per-kernel timing with HIP events on each batch's stream
    kernel_ms = {}
    for name, b in zip(("verify_set",) if combined else ("verify", "proof_verify"), sets[0]):
        _, km = b.profile(2)
        kernel_ms[name] = {k: round(v, 3) for k, v in km.items()}
        log("%s kernels (ms): %s" % (name, kernel_ms[name]))
    # the pairing kernel alone, on a Verify batch checked item by item (the
    # timed sets aggregate their checks, so their per-item launches hold only
    # the items of failing groups): 8192 full checks
    rb = g.StagedBatch(g.BATCH_VERIFY, vreqs)
    _, rk = rb.profile(2)
    rb.free()
    pairing_ms = rk.get("pairing", 0.0)
    checks = N_VERIFY
    rc, mad_ms = g.mad_peak_bench(1 << 20, 256)
    peak_mad_s = (1 << 20) * 256 * 64 / (mad_ms * 1e-3)
    step_mad_s = step_fpm * MADS_PER_FPM / (ms_per_step * 1e-3)
    pair_mad_s = checks * FPM_PER_PAIRING_CHECK * MADS_PER_FPM / (pairing_ms * 1e-3)
    units = {k: sum(w[k] for w in step_work) for k in step_work[0] if k != "items"}
    roofline = {"bound": "valu",
                "kernel": "whole step: every kernel of the timed Verify+ProofVerify sets (comb, GLV and "
                          "aggregation products, decode, reductions, pairing checks)",
                "achieved": round(step_mad_s / 1e12, 3), "peak": round(peak_mad_s / 1e12, 3),
                "unit": "Tmad/s (v_mad_u64_u32)", "frac": round(step_mad_s / peak_mad_s, 4),
                "traffic": profiled(PROFILE_STEP, "traffic_bytes_per_step"),
                "traffic_unit": "HBM bytes per step (rocprofv3 FETCH_SIZE/WRITE_SIZE of this configuration)",
                "work_per_step": "%.4g FpM x %d mads (cost_model.json) over %s" % (step_fpm, MADS_PER_FPM, units),
                "measured_int64_lane_ops_per_step": profiled(PROFILE_STEP, "int64_lane_ops_per_step"),
                "peak_source": "k_mad_peak: 2^20 lanes x 256 independent v_mad_u64_u32, HIP events",
                "pairing_kernel": {
                    "kernel": "k_pairing on a Verify batch of 8192 per-item checks (secondary)",
                    "achieved": round(pair_mad_s / 1e12, 3), "frac": round(pair_mad_s / peak_mad_s, 4),
                    "launch_ms": round(pairing_ms, 3),
                    "traffic": profiled(PROFILE_PAIRING, "traffic_bytes_per_launch"),
                    "work_per_launch": "%d checks x %d FpM x %d mads" % (checks, FPM_PER_PAIRING_CHECK,
                                                                       MADS_PER_FPM)}}
    out = {"metric": "Verify+ProofVerify ops/s on 1 MI355X; p50/p99 latency @ fixed load",
           "value": round(value, 1), "unit": "ops/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "u32 (exact Fp/Fr)",
           "data": "synthetic, seeded; keys/signatures/proofs made by the library",
           "config": {"workload": "8192 batched uneven Verify (256 issuers, L log-uniform 1..128, "
                                  "32-256 B messages, 64 B header, ~1% tampered) + 2048 ProofVerify "
                                  "(L=32 x 64 B, 25% disclosed, 64 B ph)",
                      "verify_items": N_VERIFY, "proof_verify_items": N_PROOFS,
                      "launch": "kernel by kernel" if args.no_graph else "HIP graph replay per staged batch",
                      "batch_sets_in_flight": nset,
                      "agg_straus": straus_used,
                      "comb_chunk": args.comb_chunk or (tuned or {}).get("comb_chunk", 4),
                      "staging": ("verify set: the Verify and the ProofVerify batch staged as one batch, "
                                  "every stage launched once for both" if combined
                                  else "a Verify batch and a ProofVerify batch per set"),
                      "pipeline_tuning": (dict(tuned, source=os.path.relpath(PIPELINE_TUNING, ROOT))
                                          if tuned else None),
                      "streams_per_batch": 1 if single else 2,
                      "pairing_checks": ("per-key groups of %d as random linear combinations%s, "
                                         "per-item fallback for failing groups"
                                         % (args.agg_group or (tuned["agg_group"] if tuned else 8),
                                            (", checked %d groups at a time first" % sup_used)
                                            if sup_used else "")) if agg else "per item",
                      "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                      "msm_terms": msm_terms},
           "roofline": roofline,
           "abi_calls": abi_ops,
           "kernel_ms": kernel_ms}
    for b in sets[0]:
        b.free()
    # the MSM shapes are tuned (by measurement) after the Verify/ProofVerify
    # timing, which uses fixed shapes and no tuning table
    tuning = None
    if not args.no_autotune and not args.no_msm:
        t0 = time.time()
        rc, rep = g.autotune(16)
        if rc != 0:
            raise SystemExit("autotune failed rc=%d" % rc)
        tuning = rep.strip().split("\n")
        out["tuning"] = tuning
        log("autotuned in %.1fs: %s" % (time.time() - t0, tuning))
    if not args.no_msm:
        t = time.time()
        out["msm"] = msm_run(g, steps=3 if args.profile_only else 10, warmup=1 if args.profile_only else 2,
                             graph=not args.no_graph)
        log("msm run %.1fs: %s" % (time.time() - t, {k: out["msm"][k] for k in ("msm_per_s", "ms_per_step", "comb_4096x32_ms", "pippenger_2_16_ms")}))
    if not args.profile_only:
        out["per_request_gpu_ops_s"] = round(per_request_gpu(g, vreqs), 2)
        t = time.time()
        out["abi_paths"] = abi_paths(g, vreqs, expect_v, pvreqs)
        log("reference-ABI paths %.1fs: %s" % (time.time() - t, out["abi_paths"]))
        out["proofvalue_codec"] = codec_run()
        log("proofValue codec: %s" % out["proofvalue_codec"])
        if not args.no_serving:
            t = time.time()
            out["latency"] = serving_child(args)
            log("serving run %.1fs: %s" % (time.time() - t, out["latency"]))
        if not args.no_cpu_baseline:
            out["cpu_baseline"] = cpu_baseline(vreqs, pvreqs)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
