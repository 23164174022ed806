// Serving loop: continuous batching of mixed BBS requests over a few device
// slots.
//
// Requests are copied at submit time into owned storage (the caller's memory
// is never retained) and queued.  Each slot is a worker thread with its own
// HIP streams and a reusable Arena of device and pinned host memory.  A free
// worker takes every pending request up to its share of the HBM budget
// (budget / slots, by a per-request size estimate) and max_batch, plans one
// sub-batch per operation kind, each on its own stream of the slot, launches
// each as soon as it is planned and waits for them.  Requests that arrive while other slots' batches are on the
// device are admitted by the next free slot, so batches of different slots
// overlap on the GPU, and batch size follows the load.  Issuer public keys
// (prepared G2 lines), generator comb tables and -G2 lines stay resident in
// the library context across batches.
//
// The batched whole-op paths are the same planners the C ABI entries use
// (bbs_batch.cpp), so a request's result does not depend on what it was
// batched with.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cubbs_bbs_gpu_abi.h"
#include "../../include/cubbs_hip_ext.h"
#include "bbs_batch.h"
#include "host_ctx.h"

namespace cubbs {
namespace {

using Clock = std::chrono::steady_clock;

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             Clock::now().time_since_epoch())
      .count();
}

struct Req {
  int kind = 0;
  int64_t ticket = 0;
  uint64_t t_arrive = 0, t_done = 0;
  size_t est = 0;
  // owned copies of every byte the request references
  std::vector<uint8_t> bytes;
  std::vector<CuBbsBbsGpuByteSlice> msgs;
  std::vector<size_t> idx;
  uint8_t seed[32];
  CuBbsBbsGpuVerifyItem v{};
  CuBbsBbsGpuProofVerifyItem pv{};
  CuBbsBbsGpuSignItem s{};
  CuBbsBbsGpuProofGenItem pg{};
  // result
  bool done = false;
  CuBbsBbsGpuResult res{};
  std::vector<uint8_t> out;
};

// Copies slices into one owned buffer; pointers are fixed up after all copies.
struct Packer {
  std::vector<std::pair<const uint8_t*, size_t>> parts;
  size_t add(const CuBbsBbsGpuByteSlice& s) {
    parts.emplace_back(s.data, s.len);
    return parts.size() - 1;
  }
  void finish(Req& r, std::vector<CuBbsBbsGpuByteSlice>& slices) {
    size_t total = 0;
    for (auto& p : parts) total += p.second;
    r.bytes.resize(std::max<size_t>(total, 1));
    slices.resize(parts.size());
    size_t off = 0;
    for (size_t i = 0; i < parts.size(); i++) {
      if (parts[i].second) std::memcpy(r.bytes.data() + off, parts[i].first, parts[i].second);
      slices[i] = CuBbsBbsGpuByteSlice{r.bytes.data() + off, parts[i].second};
      off += parts[i].second;
    }
  }
};

bool slice_valid(const CuBbsBbsGpuByteSlice& s) { return s.len == 0 || s.data != nullptr; }
bool msgs_valid(const CuBbsBbsGpuByteSlice* m, size_t n) {
  if (n && !m) return false;
  for (size_t i = 0; i < n; i++)
    if (!slice_valid(m[i])) return false;
  return true;
}

// Device bytes a request adds to a batch (staged inputs + work area), an
// upper estimate: payload twice (blob + job records), per message term
// about 0.4 KiB (scalar, term records, comb partial share), per item 4 KiB,
// per variable-base product its table workspace, per pairing check its line
// blocks.
size_t estimate(size_t payload, size_t terms, size_t products, size_t checks = 0) {
  return 4096 + 2 * payload + 400 * (terms + 4) + 4 * VAR_WS_WORDS * products +
         4 * PAIR_L2_WORDS * checks;
}

// Builds an owned request; returns 0 or an error code.
int32_t copy_request(const CubbsHipRequest& q, Req& r) {
  Packer pk;
  std::vector<CuBbsBbsGpuByteSlice> sl;
  r.kind = q.kind;
  if (!q.item) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  switch (q.kind) {
    case CUBBS_HIP_REQ_VERIFY: {
      const auto& it = *static_cast<const CuBbsBbsGpuVerifyItem*>(q.item);
      if (!slice_valid(it.public_key) || !slice_valid(it.signature) || !slice_valid(it.header) ||
          !msgs_valid(it.messages, it.message_count))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      pk.add(it.public_key);
      pk.add(it.signature);
      pk.add(it.header);
      for (size_t i = 0; i < it.message_count; i++) pk.add(it.messages[i]);
      pk.finish(r, sl);
      r.msgs.assign(sl.begin() + 3, sl.end());
      r.v = CuBbsBbsGpuVerifyItem{sl[0], sl[1], sl[2], r.msgs.data(), r.msgs.size()};
      r.est = estimate(r.bytes.size(), it.message_count + 2, 3, 2);
      return 0;
    }
    case CUBBS_HIP_REQ_PROOF_VERIFY: {
      const auto& it = *static_cast<const CuBbsBbsGpuProofVerifyItem*>(q.item);
      if (!slice_valid(it.public_key) || !slice_valid(it.proof) || !slice_valid(it.header) ||
          !slice_valid(it.presentation_header) ||
          !msgs_valid(it.disclosed_messages, it.disclosed_count) ||
          (it.disclosed_indexes.len && !it.disclosed_indexes.data))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      pk.add(it.public_key);
      pk.add(it.proof);
      pk.add(it.header);
      pk.add(it.presentation_header);
      for (size_t i = 0; i < it.disclosed_count; i++) pk.add(it.disclosed_messages[i]);
      pk.finish(r, sl);
      r.msgs.assign(sl.begin() + 4, sl.end());
      r.idx.assign(it.disclosed_indexes.data, it.disclosed_indexes.data + it.disclosed_indexes.len);
      r.pv = CuBbsBbsGpuProofVerifyItem{sl[0], sl[1], sl[2], sl[3], r.msgs.data(), r.msgs.size(),
                                        CuBbsBbsGpuIndexSlice{r.idx.data(), r.idx.size()},
                                        it.total_message_count};
      r.est = estimate(r.bytes.size(), it.total_message_count + 8, 6, 2);
      return 0;
    }
    case CUBBS_HIP_REQ_SIGN: {
      const auto& it = *static_cast<const CuBbsBbsGpuSignItem*>(q.item);
      if (!slice_valid(it.secret_key) || !slice_valid(it.public_key) || !slice_valid(it.header) ||
          !msgs_valid(it.messages, it.message_count))
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      pk.add(it.secret_key);
      pk.add(it.public_key);
      pk.add(it.header);
      for (size_t i = 0; i < it.message_count; i++) pk.add(it.messages[i]);
      pk.finish(r, sl);
      r.msgs.assign(sl.begin() + 3, sl.end());
      r.out.assign(80, 0);
      r.s = CuBbsBbsGpuSignItem{sl[0], sl[1], sl[2], r.msgs.data(), r.msgs.size(),
                                CuBbsBbsGpuByteBuffer{r.out.data(), r.out.size(), 0}};
      r.est = estimate(r.bytes.size(), it.message_count + 2, 0);
      return 0;
    }
    case CUBBS_HIP_REQ_PROOF_GEN: {
      const auto& it = *static_cast<const CuBbsBbsGpuProofGenItem*>(q.item);
      if (!slice_valid(it.public_key) || !slice_valid(it.signature) || !slice_valid(it.header) ||
          !slice_valid(it.presentation_header) || !msgs_valid(it.messages, it.message_count) ||
          (it.disclosed_indexes.len && !it.disclosed_indexes.data) ||
          it.disclosed_indexes.len > it.message_count)
        return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
      pk.add(it.public_key);
      pk.add(it.signature);
      pk.add(it.header);
      pk.add(it.presentation_header);
      for (size_t i = 0; i < it.message_count; i++) pk.add(it.messages[i]);
      pk.finish(r, sl);
      r.msgs.assign(sl.begin() + 4, sl.end());
      r.idx.assign(it.disclosed_indexes.data, it.disclosed_indexes.data + it.disclosed_indexes.len);
      const size_t U = it.message_count - it.disclosed_indexes.len;
      r.out.assign(272 + 32 * U, 0);
      r.pg = CuBbsBbsGpuProofGenItem{sl[0], sl[1], sl[2], sl[3], r.msgs.data(), r.msgs.size(),
                                     CuBbsBbsGpuIndexSlice{r.idx.data(), r.idx.size()},
                                     CuBbsBbsGpuByteBuffer{r.out.data(), r.out.size(), 0}};
      if (q.seed32)
        std::memcpy(r.seed, q.seed32, 32);
      else
        fill_os_random(r.seed, 32);
      r.est = estimate(r.bytes.size(), 2 * it.message_count + 12, 7);
      return 0;
    }
    default:
      return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  }
}

struct SlotGraphs;

// Device bytes of the library's resident state (tables, keys, retired buffers).
size_t resident_bytes() {
  Context& c = ctx();
  std::lock_guard<std::mutex> g(c.mu);
  return c.resident_bytes;
}

struct Server {
  CubbsHipServerConfig cfg{};
  // a slot's share of the budget once the resident state is counted
  size_t slot_share() const {
    const size_t res = resident_bytes();
    return cfg.hbm_budget_bytes > res ? (size_t)((cfg.hbm_budget_bytes - res) / cfg.slots) : 0;
  }
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_space;
  std::deque<Req*> pending;
  // one free slot at a time collects a batch (and waits out the coalescing
  // window), so that slots do not take each other's partial batches
  bool collecting = false;
  // batcher statistics (under mu): summed queue wait of dispatched requests,
  // the largest and the last batch, batches per size (index min(size, cap))
  uint64_t wait_ns = 0, max_batch_seen = 0, last_batch = 0;
  std::vector<uint64_t> size_hist;
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  int64_t next_ticket = 0;
  bool stop = false;
  std::vector<std::thread> workers;
  // statistics
  std::atomic<uint64_t> batches{0}, requests{0}, on_device{0}, max_on_device{0}, peak_bytes{0};
  // summed over batches: host time planning + packing the sub-batches, and
  // the whole batch from the first plan to the last result
  std::atomic<uint64_t> plan_ns{0}, batch_ns{0};

  std::atomic<int32_t> fault{0};
  std::atomic<uint64_t> failed_batches{0};  // batches failed alone (memory, host errors)

  void worker(unsigned slot);
  void run_batch(std::vector<Req*>& take, Arena& ar, SlotGraphs* sg);
};

// The instantiated graph of each sub-batch kind of a slot, and the private
// streams its launch sequences are captured on.
struct SlotGraphs {
  hipStream_t cap[4] = {}, cap_aux[4] = {};
  hipGraphExec_t exec[4] = {};
  ~SlotGraphs() {
    for (int k = 0; k < 4; k++) {
      if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
      if (cap[k]) (void)hipStreamDestroy(cap[k]);
      if (cap_aux[k]) (void)hipStreamDestroy(cap_aux[k]);
    }
  }
};

// Launch a planned sub-batch: captured and replayed through the slot's graph
// of its kind, or kernel by kernel.
void launch_sub(Batch& b, SlotGraphs* sg, int k) {
  if (!sg) {
    launch_batch(b);
    return;
  }
  if (!sg->cap[k]) {
    check(hipStreamCreateWithFlags(&sg->cap[k], hipStreamNonBlocking));
    check(hipStreamCreateWithFlags(&sg->cap_aux[k], hipStreamNonBlocking));
  }
  const hipStream_t s0 = b.stream, a0 = b.aux;
  b.stream = sg->cap[k];
  b.aux = a0 ? sg->cap_aux[k] : nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamBeginCapture(b.stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    try {
      launch_batch(b);
    } catch (const GpuError&) {
      e = hipErrorUnknown;
    }
    const hipError_t e2 = hipStreamEndCapture(b.stream, &g);
    if (e == hipSuccess) e = e2;
  }
  b.stream = s0;
  b.aux = a0;
  if (e == hipSuccess && sg->exec[k]) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult r{};
    if (hipGraphExecUpdate(sg->exec[k], g, &bad, &r) != hipSuccess || r != hipGraphExecUpdateSuccess) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(sg->exec[k]);
      sg->exec[k] = nullptr;
    }
  }
  if (e == hipSuccess && !sg->exec[k]) e = hipGraphInstantiate(&sg->exec[k], g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  check(e);
  check(hipGraphLaunch(sg->exec[k], b.stream));
}

void Server::run_batch(std::vector<Req*>& take, Arena& ar, SlotGraphs* sg) {
  std::vector<Req*> by[4];
  for (Req* r : take) by[r->kind].push_back(r);
  std::vector<std::unique_ptr<Batch>> bs(4);
  std::vector<CuBbsBbsGpuVerifyItem> vi;
  std::vector<CuBbsBbsGpuProofVerifyItem> pvi;
  std::vector<CuBbsBbsGpuSignItem> si;
  std::vector<CuBbsBbsGpuProofGenItem> pgi;
  std::vector<uint8_t> seeds;
  struct OnDevice {  // counts this batch as on the device until it leaves, by any path
    std::atomic<uint64_t>& n;
    ~OnDevice() {
      --n;
      set_thread_arena(nullptr);
      set_thread_agg_group(0);
    }
  } guard{on_device};
  const uint64_t od = ++on_device;
  uint64_t m = max_on_device.load();
  while (od > m && !max_on_device.compare_exchange_weak(m, od)) {
  }
  // each kind is planned onto its own stream of the slot and launched as soon
  // as it is planned, so the host packs the next kind while the device runs
  const uint64_t t0 = now_ns();
  uint64_t t_plan = 0;
  auto planned = [&](uint64_t ts) { t_plan += now_ns() - ts; };
  set_thread_arena(&ar);
  set_thread_agg_group((cfg.flags & CUBBS_HIP_SERVER_AGGREGATE) ? 8u : 0u);
  if (!by[CUBBS_HIP_REQ_VERIFY].empty()) {
    for (Req* r : by[CUBBS_HIP_REQ_VERIFY]) vi.push_back(r->v);
    ar.cur = 0;
    const uint64_t ts = now_ns();
    bs[0] = plan_verify(vi.data(), vi.size());
    planned(ts);
    launch_sub(*bs[0], sg, 0);
  }
  if (!by[CUBBS_HIP_REQ_PROOF_VERIFY].empty()) {
    for (Req* r : by[CUBBS_HIP_REQ_PROOF_VERIFY]) pvi.push_back(r->pv);
    ar.cur = 1;
    const uint64_t ts = now_ns();
    bs[1] = plan_proof_verify(pvi.data(), pvi.size());
    planned(ts);
    launch_sub(*bs[1], sg, 1);
  }
  set_thread_agg_group(0);
  if (!by[CUBBS_HIP_REQ_SIGN].empty()) {
    for (Req* r : by[CUBBS_HIP_REQ_SIGN]) si.push_back(r->s);
    ar.cur = 2;
    const uint64_t ts = now_ns();
    bs[2] = plan_sign(si.data(), si.size());
    planned(ts);
    launch_sub(*bs[2], sg, 2);
  }
  if (!by[CUBBS_HIP_REQ_PROOF_GEN].empty()) {
    for (Req* r : by[CUBBS_HIP_REQ_PROOF_GEN]) {
      pgi.push_back(r->pg);
      seeds.insert(seeds.end(), r->seed, r->seed + 32);
    }
    ar.cur = 3;
    const uint64_t ts = now_ns();
    bs[3] = plan_proof_gen(pgi.data(), pgi.size(), 1, seeds.data(), nullptr, nullptr);
    planned(ts);
    launch_sub(*bs[3], sg, 3);
  }
  set_thread_arena(nullptr);
  const uint64_t held = ar.dev_bytes;
  uint64_t pb = peak_bytes.load();
  while (held > pb && !peak_bytes.compare_exchange_weak(pb, held)) {
  }
  for (int k = 0; k < 4; k++) {
    if (!bs[k]) continue;
    std::vector<CuBbsBbsGpuResult> res(by[k].size());
    std::vector<uint8_t*> outs(by[k].size());
    for (size_t i = 0; i < by[k].size(); i++) outs[i] = by[k][i]->out.data();
    collect_batch(*bs[k], res.data(), outs.data(), nullptr);
    for (size_t i = 0; i < by[k].size(); i++) by[k][i]->res = res[i];
  }
  bs.clear();
  ar.reset();
  plan_ns += t_plan;
  batch_ns += now_ns() - t0;
}

void Server::worker(unsigned slot) {
  (void)slot;
  Arena ar;
  SlotGraphs graphs;
  SlotGraphs* sg = (cfg.flags & CUBBS_HIP_SERVER_GRAPHS) ? &graphs : nullptr;
  try {
    {
      Context& c = ctx();
      std::lock_guard<std::mutex> g(c.mu);
      init_locked(c);
      for (int i = 0; i < 4; i++) {
        ar.streams[i] = next_stream_locked(c);
        ar.aux[i] = next_stream_locked(c);
      }
    }
    for (int i = 0; i < 4; i++) {
      check(hipEventCreateWithFlags(&ar.fork[i], hipEventDisableTiming));
      check(hipEventCreateWithFlags(&ar.join[i], hipEventDisableTiming));
    }
  } catch (...) {
    fault = CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  for (;;) {
    std::vector<Req*> take;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] { return stop || (!pending.empty() && !collecting); });
      if (pending.empty() && stop) return;
      if (pending.empty()) continue;
      // a partial batch waits up to coalesce_us for more requests (the
      // reference batcher's interval: a window for a partial batch, not a
      // throttle; a full batch goes at once)
      if (cfg.coalesce_us && pending.size() < cfg.max_batch && !stop) {
        collecting = true;
        const auto deadline = Clock::now() + std::chrono::microseconds(cfg.coalesce_us);
        cv_work.wait_until(lk, deadline, [&] { return stop || pending.size() >= cfg.max_batch; });
        collecting = false;
      }
      size_t bytes = 0;
      const size_t share = slot_share();
      while (!pending.empty() && take.size() < cfg.max_batch) {
        Req* r = pending.front();
        if (!take.empty() && bytes + r->est > share) break;
        bytes += r->est;
        take.push_back(r);
        pending.pop_front();
      }
      if (!take.empty()) {
        const uint64_t t = now_ns();
        for (const Req* r : take) wait_ns += t > r->t_arrive ? t - r->t_arrive : 0;
        max_batch_seen = std::max<uint64_t>(max_batch_seen, take.size());
        last_batch = take.size();
        size_hist[std::min<size_t>(take.size(), size_hist.size() - 1)]++;
      }
    }
    // the queue has room again, and another free slot may collect the rest
    cv_space.notify_all();
    cv_work.notify_all();
    if (take.empty()) continue;
    // A batch that fails for want of memory (device or host) or on a host
    // error fails alone: its requests get code 3 and the slot goes on.  Any
    // other device error leaves the device in an unknown state: the server
    // then fails every later request too (the fault is reported by stats).
    bool failed = false;
    if (fault == 0) {
      try {
        run_batch(take, ar, sg);
      } catch (const GpuError& e) {
        failed = true;
        if (e.err != hipErrorOutOfMemory) fault = CUBBS_BBS_GPU_ERR_INTERNAL;
      } catch (...) {
        failed = true;
      }
      if (failed) {
        // the batches of the slot may still run: wait for every stream of the
        // slot before its memory is used again
        for (int i = 0; i < 4; i++) {
          if (ar.streams[i]) (void)hipStreamSynchronize(ar.streams[i]);
          if (ar.aux[i]) (void)hipStreamSynchronize(ar.aux[i]);
        }
        for (int k = 0; k < 4; k++) {
          if (graphs.cap[k]) (void)hipStreamSynchronize(graphs.cap[k]);
          if (graphs.exec[k]) (void)hipGraphExecDestroy(graphs.exec[k]);
          graphs.exec[k] = nullptr;
        }
        (void)hipGetLastError();
        ar.reset();
        failed_batches++;
      }
    }
    const uint64_t t = now_ns();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Req* r : take) {
        if (fault || failed) {
          r->res = CuBbsBbsGpuResult{CUBBS_BBS_GPU_ERR_INTERNAL, 0, 0};
        }
        r->t_done = t;
        r->done = true;
      }
      batches++;
      requests += take.size();
    }
    cv_done.notify_all();
  }
}

int64_t submit_at(Server* s, const CubbsHipRequest* q, uint64_t t_arrive) {
  auto r = std::make_unique<Req>();
  int32_t rc = copy_request(*q, *r);
  if (rc) return -(int64_t)rc;
  // a request above a slot's share runs in a batch of its own; one above the
  // whole HBM budget is refused at admission (code 1) instead of failing on
  // the device
  if (r->est > s->cfg.hbm_budget_bytes) return -(int64_t)CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  r->t_arrive = t_arrive;
  std::unique_lock<std::mutex> lk(s->mu);
  // a full queue blocks the submitter until a slot takes requests (the
  // reference batcher's bounded channel)
  if (s->cfg.queue_capacity)
    s->cv_space.wait(lk, [&] { return s->stop || s->pending.size() < s->cfg.queue_capacity; });
  if (s->stop) return -(int64_t)CUBBS_BBS_GPU_ERR_INTERNAL;
  r->ticket = s->next_ticket++;
  Req* raw = r.get();
  s->reqs.emplace(raw->ticket, std::move(r));
  s->pending.push_back(raw);
  // every worker: a collecting slot waits for its batch to fill
  s->cv_work.notify_all();
  return raw->ticket;
}

int32_t wait_ticket(Server* s, int64_t ticket, CuBbsBbsGpuResult* result, uint8_t* out,
                    size_t out_cap, uint64_t* latency_ns) {
  std::unique_ptr<Req> r;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
    Req* raw = it->second.get();
    s->cv_done.wait(lk, [&] { return raw->done; });
    r = std::move(it->second);
    s->reqs.erase(it);
  }
  if (result) *result = r->res;
  if (latency_ns) *latency_ns = r->t_done - r->t_arrive;
  const size_t n = r->res.bytes_written;
  if (n && out) {
    if (out_cap < n) return CUBBS_BBS_GPU_ERR_INSUFFICIENT_OUTPUT;
    std::memcpy(out, r->out.data(), n);
  }
  return CUBBS_BBS_GPU_SUCCESS;
}

}  // namespace
}  // namespace cubbs

using cubbs::Server;

extern "C" {

void* cubbs_hip_server_create(const CubbsHipServerConfig* cfg, int32_t* status_out) {
  auto set = [&](int32_t v) {
    if (status_out) *status_out = v;
  };
  try {
    {
      cubbs::Context& c = cubbs::ctx();
      std::lock_guard<std::mutex> g(c.mu);
      cubbs::init_locked(c);
    }
    if (cfg && (cfg->flags & ~(CUBBS_HIP_SERVER_GRAPHS | CUBBS_HIP_SERVER_AGGREGATE))) {
      set(CUBBS_BBS_GPU_ERR_INVALID_INPUT);
      return nullptr;
    }
    auto* s = new Server();
    // default slots: as many as the stream pool serves without two slots
    // sharing a stream (a slot uses 8), since a launch on a stream that is
    // busy with another slot's work can stall the host thread
    uint32_t pool = 0;
    {
      cubbs::Context& c = cubbs::ctx();
      std::lock_guard<std::mutex> g(c.mu);
      pool = (uint32_t)c.pool.size();
    }
    CubbsHipServerConfig d{std::max<uint32_t>(1, pool / 8), 4096, cubbs::resident_bytes() + (4ull << 30), 0, 0};
    if (cfg) {
      if (cfg->slots) d.slots = cfg->slots;
      if (cfg->max_batch) d.max_batch = cfg->max_batch;
      if (cfg->hbm_budget_bytes) d.hbm_budget_bytes = cfg->hbm_budget_bytes;
      d.flags = cfg->flags;
    }
    d.slots = std::min<uint32_t>(d.slots, 16);
    if (cfg) {
      d.coalesce_us = std::min<uint32_t>(cfg->coalesce_us, 1000000);  // at most 1 s
      d.queue_capacity = cfg->queue_capacity;
    }
    s->cfg = d;
    s->size_hist.assign(std::min<size_t>(d.max_batch, 1u << 16) + 1, 0);
    for (unsigned i = 0; i < d.slots; i++) s->workers.emplace_back([s, i] { s->worker(i); });
    set(CUBBS_BBS_GPU_SUCCESS);
    return s;
  } catch (...) {
    set(CUBBS_BBS_GPU_ERR_INTERNAL);
    return nullptr;
  }
}

int64_t cubbs_hip_server_submit(void* srv, const CubbsHipRequest* req) {
  if (!srv || !req) return -(int64_t)CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  try {
    return cubbs::submit_at(static_cast<Server*>(srv), req, cubbs::now_ns());
  } catch (...) {  // host memory exhausted while copying the request
    return -(int64_t)CUBBS_BBS_GPU_ERR_INTERNAL;
  }
}

int32_t cubbs_hip_server_wait(void* srv, int64_t ticket, CuBbsBbsGpuResult* result, uint8_t* out,
                              size_t out_cap, uint64_t* latency_ns) {
  if (!srv) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  return cubbs::wait_ticket(static_cast<Server*>(srv), ticket, result, out, out_cap, latency_ns);
}

int32_t cubbs_hip_server_run_open_loop(void* srv, const CubbsHipRequest* reqs,
                                       const uint64_t* arrival_ns, size_t n,
                                       CuBbsBbsGpuResult* results, uint64_t* latency_ns,
                                       uint8_t* const* outs, const size_t* out_caps) {
  if (!srv || (n && (!reqs || !arrival_ns || !results || !latency_ns)))
    return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  // arrivals must be in order: checked before anything is submitted
  for (size_t i = 1; i < n; i++)
    if (arrival_ns[i] < arrival_ns[i - 1]) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  auto* s = static_cast<Server*>(srv);
  std::vector<int64_t> tickets;
  try {
    tickets.resize(n);
  } catch (...) {
    return CUBBS_BBS_GPU_ERR_INTERNAL;
  }
  const uint64_t t0 = cubbs::now_ns();
  for (size_t i = 0; i < n; i++) {
    const uint64_t due = t0 + arrival_ns[i];
    uint64_t t = cubbs::now_ns();
    if (t < due) {
      if (due - t > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(due - t - 100000));
      while (cubbs::now_ns() < due) {
      }
    }
    // latency counts from the scheduled arrival, so a late submission is charged
    try {
      tickets[i] = cubbs::submit_at(s, &reqs[i], due);
    } catch (...) {  // host memory exhausted: this request fails alone
      tickets[i] = -(int64_t)CUBBS_BBS_GPU_ERR_INTERNAL;
    }
  }
  int32_t rc = CUBBS_BBS_GPU_SUCCESS;
  for (size_t i = 0; i < n; i++) {
    if (tickets[i] < 0) {
      results[i] = CuBbsBbsGpuResult{(int32_t)-tickets[i], 0, 0};
      latency_ns[i] = 0;
      continue;
    }
    int32_t w = cubbs::wait_ticket(s, tickets[i], &results[i], outs ? outs[i] : nullptr,
                                   out_caps ? out_caps[i] : 0, &latency_ns[i]);
    if (w && !rc) rc = w;
  }
  if (s->fault) rc = s->fault;
  return rc;
}

int32_t cubbs_hip_server_stats(void* srv, uint64_t* batches, uint64_t* requests,
                               uint64_t* max_concurrent_batches, uint64_t* peak_slot_bytes) {
  if (!srv) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  auto* s = static_cast<Server*>(srv);
  if (batches) *batches = s->batches;
  if (requests) *requests = s->requests;
  if (max_concurrent_batches) *max_concurrent_batches = s->max_on_device;
  if (peak_slot_bytes) *peak_slot_bytes = s->peak_bytes;
  return s->fault;
}

int32_t cubbs_hip_server_queue_stats(void* srv, uint64_t* queued, uint64_t* total_wait_ns,
                                     uint64_t* max_batch, uint64_t* last_batch, uint64_t* hist,
                                     size_t hist_len) {
  if (!srv || (hist && !hist_len)) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  auto* s = static_cast<Server*>(srv);
  std::lock_guard<std::mutex> lk(s->mu);
  if (queued) *queued = s->pending.size();
  if (total_wait_ns) *total_wait_ns = s->wait_ns;
  if (max_batch) *max_batch = s->max_batch_seen;
  if (last_batch) *last_batch = s->last_batch;
  if (hist) {
    std::fill(hist, hist + hist_len, 0);
    for (size_t k = 0; k < s->size_hist.size(); k++) hist[std::min(k, hist_len - 1)] += s->size_hist[k];
  }
  return s->fault;
}

int32_t cubbs_hip_server_timing(void* srv, double* plan_ms_per_batch, double* batch_ms_per_batch) {
  if (!srv) return CUBBS_BBS_GPU_ERR_INVALID_INPUT;
  auto* s = static_cast<Server*>(srv);
  const double n = std::max<double>(1.0, (double)s->batches.load());
  if (plan_ms_per_batch) *plan_ms_per_batch = (double)s->plan_ns.load() / n * 1e-6;
  if (batch_ms_per_batch) *batch_ms_per_batch = (double)s->batch_ns.load() / n * 1e-6;
  return s->fault;
}

void cubbs_hip_server_destroy(void* srv) {
  if (!srv) return;
  auto* s = static_cast<Server*>(srv);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_space.notify_all();
  for (auto& t : s->workers) t.join();
  delete s;
}

}  // extern "C"
