// Serving loop: continuous batching of focusing requests against device-resident datasets.
//
// One scheduling thread per server.  Each round it admits queued requests (FIFO) while a slot and
// device memory within the budget are free, then focuses every admitted request in ONE launch of
// the persistent batch kernel (bp_batch.hip), each tile advanced by the round's pulse span.  Partial
// per-pixel sums stay in device memory between rounds; completed requests are copied out and their
// memory released, preempted ones have their partial state saved on the host.  A request with too
// few tiles to fill the device is instead focused whole in its first round, each tile's pulses
// split into segments run by separate wavefronts, whose partial sums a second kernel adds in
// segment order.  Results do not depend on the mix or on the round boundaries: the per-pixel
// arithmetic and its order are fixed by the request alone.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "batch_internal.h"
#include "serve_policy.h"

namespace {

using Clock = std::chrono::steady_clock;

// Page-locked host blocks for the requests' pixel copies and results, reused across requests:
// fresh memory per request costs page faults on the submit path, and uploads from pageable memory
// are staged through the runtime's own buffers.  Coherent (fine-grained) memory: the kernels'
// stores of the results reach host memory directly, visible once the round's event has completed.
// Same size classes and reuse rule as DevicePool; its own lock (submit runs in the callers'
// threads, finish in the scheduling and completion threads).
// The submission pass over a large request's pixels runs on up to this many threads, about 2^16
// pixels each.
constexpr int SUBMIT_MAX_THREADS = 8;

struct PinnedPool {
  std::mutex mu;
  std::multimap<size_t, char *> blocks;
  size_t cached = 0;
  size_t limit = (size_t)4 << 30;
  char *get(size_t bytes, size_t *cap) {
    const size_t want = (bytes + ((size_t)1 << 20) - 1) >> 20 << 20;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = blocks.lower_bound(want);
      if (it != blocks.end() && it->first <= 2 * want) {
        char *p = it->second;
        *cap = it->first;
        cached -= it->first;
        blocks.erase(it);
        return p;
      }
    }
    char *p = nullptr;
    if (hipHostMalloc((void **)&p, want, hipHostMallocCoherent) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    *cap = want;
    return p;
  }
  void put(char *p, size_t cap) {
    if (!p) return;
    {
      std::lock_guard<std::mutex> lk(mu);
      if (cached + cap <= limit) {
        blocks.emplace(cap, p);
        cached += cap;
        return;
      }
    }
    (void)hipHostFree(p);
  }
  void release() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto &b : blocks) (void)hipHostFree(b.second);
    blocks.clear();
    cached = 0;
  }
};

struct ServerReq {
  int64_t id = 0;
  sarbp_dataset *ds = nullptr;
  char *host = nullptr;  // page-locked copy of the pixel arrays (from the server's PinnedPool)
  size_t host_cap = 0;
  sarbp_request req{};  // points into the host copy above
  double *out = nullptr;
  int32_t *ncnt = nullptr;
  int32_t state = SARBP_REQ_QUEUED;
  int32_t precision = SARBP_PREC_F64;
  double est = 0.0;
  float r0min = 0.f;  // the pixels' distance bound (split requests; from the submission pass)
  bool narrow = false;  // FP64 terms may take the narrow-range form (from the submission pass)
  ReqBuf buf;
  bool on_device = false;
  int off = 0;       // pulses of every tile's union already focused (multiple of the chunk)
  int max_len = 0;   // longest tile union (pulses)
  int64_t need = 0;  // device bytes while admitted
  // a preempted request's device contents, saved on the host: every pixel's partial state, and the
  // means and counts already written by the tiles that finished before the preemption
  std::vector<BpState> saved;
  std::vector<int> tile_len;   // per-tile union lengths, computed at submission (caller's thread)
  bool preempt = false;
  bool completing = false;  // results on the host, being copied into the caller's buffers
  bool last_round = false;  // scheduled in a round that completes it
  size_t out_off = 0;       // the means and counts in the page-locked block, which the kernels write
  bool holds_ds = false;  // counted in ds->server_refs until done or failed
  double submit_s = 0, admit_s = -1, done_s = -1, round_s = -1, gpu_s = -1;
  int32_t rounds = 0, preemptions = 0;
  int32_t waiters = 0;  // threads in sarbp_server_wait on this request (the record must outlive them)
  int64_t pp = 0;
  int nseg = 1;     // pulse segments per tile (> 1: focused whole in its first round)
  int seg_len = 0;  // pulses per segment (a multiple of the chunk)
};

// A-priori RMS phase error (rad) of a request's split result against its FP64 result, over the
// pixels with |a| > 1e-5 max|a|, from its tile rows' geometry (SplitRows).  Per-term model: the
// FP32 error of the range offset dR ~ 1.2e-7 x |a|, with |a| = |u.o| <= o_x c + o_y + o_z, plus
// ~2e-6 rad for the base phase and the hardware sin/cos.  The sums' errors are partly correlated
// along the aperture and dark pixels weigh in the RMS, so the measured RMS error is 0.2-2.0 times
// this model (72 requests on grids 1-6 m apart, Sint 400-3500); the estimate is the model times 3.
// The range series' truncation adds kabs x split_truncation_m (a systematic per-term error).
double split_phase_estimate(const sarbp_dataset *ds, const SplitRows &rows) {
  if (!rows.any) return 0.0;
  const double kabs = std::fabs(-4.0 * SARBP_PI32_F64 / ds->lambda_val);
  const double trunc = kabs * split_truncation_m(rows.omax, rows.r0min);
  // rows this wide, or this close to the track, take FP64 terms in the kernel anyway
  if (rows.omax > (double)SPLIT_OMAX_M || trunc > SPLIT_TRUNC_MAX_RAD) return 1e9;
  return 3.0 * (kabs * 1.2e-7 * rows.amax + 2e-6) + trunc;
}

}  // namespace

// One round of the serving loop: its requests, those it completes, its work-list buffers and its
// events.  Two alternate, so that the next round is built and queued while one runs.
struct Round {
  std::vector<ServerReq *> reqs, fin;
  int64_t pp = 0;  // pixel-pulse pairs of its work items
  double t_round = 0.0;
  int max_rows = 1;
  int waves = 1;  // wavefronts per block of its batch launch = adjacent tiles per work item
  bool any_f64 = false, any_split = false;
  bool live = false;    // launched, not yet collected
  bool failed = false;  // a launch or copy of it failed
  bool timed = false;   // e0 and e1 bracket its kernels
  ItemBuffers ib;
  hipEvent_t e0 = nullptr, e1 = nullptr;  // around its batch kernel (and the slot clear before it)
  hipEvent_t ek = nullptr;                 // after its batch kernel
  hipEvent_t done = nullptr;               // after its segments' combine: every result written
};

struct sarbp_server {
  sarbp_server_config cfg{};
  std::thread th;
  // completion threads: copy finished requests' results from their page-locked blocks into the
  // callers' buffers, so that the scheduling thread goes on to the next round at once
  static constexpr int COMPLETION_THREADS = 2;
  std::thread th_done[COMPLETION_THREADS];
  std::deque<ServerReq *> completing;
  std::condition_variable cv_complete;
  bool stop_complete = false;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int64_t, std::unique_ptr<ServerReq>> reqs;
  std::deque<ServerReq *> queue;
  std::vector<ServerReq *> active;
  int64_t next_id = 1;
  bool stop = false;
  bool started = false;
  int device = 0;
  int64_t slots_f64 = 4096, slots_split = 4096;  // resident wavefronts of the FP64 / split batch kernels
  // three streams at the default priority: rounds' kernels (st), admissions' uploads (st_up), so
  // that the uploads of the next round's requests overlap this round's kernels, and the segmented
  // requests' combine (st_comb).  Nothing but copies goes on st_up: a kernel there would wait for
  // the persistent round kernel to leave the CUs, and the copies queued behind it with it.  The
  // results need no copy: the kernels write them into the requests' page-locked blocks; the
  // combine, whose time is those writes over the host link, runs beside the next round's batch
  // kernel (it needs few registers and no LDS, so its blocks fit beside the persistent blocks).
  hipStream_t st = nullptr, st_up = nullptr, st_comb = nullptr;
  hipEvent_t ev_up = nullptr;  // after the uploads of the requests admitted last
  bool up_pending = false;     // the next round waits for ev_up
  Round rounds[2];
  std::vector<BatchReqDev> devs;  // the round being built: its work list (scheduling thread only, reused)
  std::vector<int32_t> segmented;  // its requests with pulse segments (indices into devs)
  std::vector<WorkItemCost> work;
  std::vector<BatchItem> items;
  std::vector<int32_t> seg;
  DevicePool pool;    // request blocks (scheduling thread only)
  PinnedPool hpool;   // host copies of the requests' pixels
  int64_t bytes_in_use = 0;
  double rate = 3e11;  // pixel-pulse pairs per second, running estimate
  Clock::time_point t0 = Clock::now();
  sarbp_server_stats stats{};

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  void sync_all() {
    for (hipStream_t q : {st, st_up, st_comb})
      if (q) (void)hipStreamSynchronize(q);
  }
  void destroy_streams() {
    for (hipStream_t *q : {&st, &st_up, &st_comb})
      if (*q) {
        (void)hipStreamDestroy(*q);
        *q = nullptr;
      }
  }
  void destroy_events() {
    if (ev_up) (void)hipEventDestroy(ev_up);
    ev_up = nullptr;
    for (Round &R : rounds)
      for (hipEvent_t *ev : {&R.e0, &R.e1, &R.ek, &R.done})
        if (*ev) {
          (void)hipEventDestroy(*ev);
          *ev = nullptr;
        }
  }
  void run();
  bool build_round(Round &R);  // under mu
  void launch_round(Round &R);  // without mu
  void collect_round(Round &R, std::unique_lock<std::mutex> &lk);
  void fail_rounds(std::unique_lock<std::mutex> &lk);
  void complete_loop();
  bool admit(ServerReq *r);
  void release(ServerReq *r) {
    reqbuf_free(r->buf);
    r->on_device = false;
    bytes_in_use -= r->need;
  }
  static void drop_saved(ServerReq *r) {
    std::vector<BpState>().swap(r->saved);
  }
  // A finished (done or failed) request keeps only its record: its host copy of the pixel arrays
  // and any saved state are dropped, so a long-running server holds the pixels of the requests in
  // flight only.
  static void drop_ds(ServerReq *r) {
    if (r->holds_ds) r->ds->server_refs.fetch_sub(1);
    r->holds_ds = false;
  }
  void finish(ServerReq *r, int32_t state) {
    r->state = state;
    drop_ds(r);
    hpool.put(r->host, r->host_cap);  // no upload from it is pending: the request left the device
    r->host = nullptr;
    r->host_cap = 0;
    drop_saved(r);
    std::vector<int>().swap(r->tile_len);
    r->req.x = r->req.y = r->req.z = nullptr;
    r->req.mask = nullptr;
    r->req.index = r->req.sint = nullptr;
  }
};

void sarbp_server::complete_loop() {
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    cv_complete.wait(lk, [this] { return stop_complete || !completing.empty(); });
    if (completing.empty()) break;  // stop_complete and nothing left
    ServerReq *r = completing.front();
    completing.pop_front();
    lk.unlock();
    if (r->buf.n > 0 && r->host) {  // the means at out_off, the counts after them (256-byte aligned)
      const size_t n = (size_t)r->buf.n;
      if (r->out) std::memcpy(r->out, r->host + r->out_off, n * sizeof(double2));
      if (r->ncnt) std::memcpy(r->ncnt, r->host + r->out_off + (n * sizeof(double2) + 255) / 256 * 256, n * sizeof(int32_t));
    }
    lk.lock();
    char *blk = r->host;  // back to the pool after the lock is released: a block beyond the pool's
    const size_t cap = r->host_cap;  // limit is freed, and freeing waits for the device
    r->host = nullptr;
    r->host_cap = 0;
    r->completing = false;
    r->done_s = now();
    finish(r, SARBP_REQ_DONE);
    stats.completed++;
    stats.pp_done += r->pp;
    cv_done.notify_all();
    lk.unlock();
    hpool.put(blk, cap);
    lk.lock();
  }
}

bool sarbp_server::admit(ServerReq *r) {
  // the uploads are ordered before the round's launch on the same stream; only a saved state,
  // whose host copy is dropped below, needs the stream drained first
  int rc = reqbuf_create(r->ds, r->req, r->nseg, default_tiling(r->precision).accum, r->buf, st_up, &pool, &r->tile_len,
                         r->r0min, r->host, r->host ? r->host + r->out_off : nullptr, r->narrow ? 1 : 0);
  if (rc) return false;
  if (r->off == 0) drop_saved(r);  // no progress to restore
  if (!r->saved.empty()) {
    // the partial sums of unfinished tiles (the means of the tiles that finished before the
    // preemption are in the request's page-locked block, which the kernels write)
    const size_t n = r->saved.size();
    hipError_t e = hipMemcpyAsync(r->buf.state, r->saved.data(), n * sizeof(BpState), hipMemcpyHostToDevice, st_up);
    if (e == hipSuccess) e = hipStreamSynchronize(st_up);
    if (e != hipSuccess) {
      (void)hipStreamSynchronize(st_up);
      reqbuf_free(r->buf);
      return false;
    }
  }
  drop_saved(r);
  r->on_device = true;
  r->max_len = 0;
  for (int v : r->buf.tile_len) r->max_len = std::max(r->max_len, v);
  return true;
}

bool sarbp_server::build_round(Round &R) {
  const int chunk = cfg.pulse_chunk;
  R.reqs.clear();
  R.fin.clear();
  for (ServerReq *r : active)
    if (!r->last_round) R.reqs.push_back(r);
  if (R.reqs.empty()) return false;
  // every request of the round advances every unfinished tile by `span` pulses
  int64_t live_px = 0;
  for (ServerReq *r : R.reqs)
    if (r->nseg == 1) live_px += std::max<int64_t>(1, r->buf.n);
  const double budget_pp = rate * cfg.quantum_ms * 1e-3;
  int span = (int)std::min<double>(1 << 30, std::ceil(budget_pp / (double)std::max<int64_t>(1, live_px)));
  span = std::max(chunk, (span + chunk - 1) / chunk * chunk);
  const double t_round = now();
  R.t_round = t_round;
  devs.clear();
  work.clear();
  segmented.clear();
  R.max_rows = 1;
  R.pp = 0;
  R.any_f64 = R.any_split = false;
  for (ServerReq *r : R.reqs) {
    (r->precision == SARBP_PREC_SPLIT ? R.any_split : R.any_f64) = true;
    if (r->precision == SARBP_PREC_SPLIT) R.max_rows = std::max(R.max_rows, r->buf.tg.ty);
  }
  // four wavefronts per block share the block's sin/cos table and take four adjacent tiles per work
  // item (measured on the 64-request batch: 8 % faster than one), unless the records of long pulse
  // chunks do not fit the LDS
  R.waves = 4;
  while (R.waves > 1 && sarbp_batch_lds_bytes(R.waves, chunk, R.max_rows) > SARBP_BATCH_LDS_MAX) R.waves >>= 1;
  const int W = R.waves;
  for (size_t k = 0; k < R.reqs.size(); ++k) {
    ServerReq *r = R.reqs[k];
    r->round_s = t_round;
    devs.push_back(reqbuf_dev(r->ds, r->buf));
    devs.back().nseg = r->nseg;
    if (r->nseg > 1 && r->buf.n > 0) segmented.push_back((int32_t)k);
    const int64_t nt = r->buf.tg.ntiles();
    // one item per W adjacent tiles and pulse range; a tile of the group with nothing left in the
    // range returns at once in the kernel (its cost is the group's longest)
    for (int64_t t0 = 0; t0 < nt; t0 += W) {
      const int64_t t1 = std::min<int64_t>(nt, t0 + W);
      if (r->nseg > 1) {  // every segment of every tile, this round
        for (int j = 0; j < r->nseg; ++j) {
          const int a = j * r->seg_len;
          int cmax = 0;
          for (int64_t t = t0; t < t1; ++t) {
            const int len = r->buf.tile_len[(size_t)t];
            if (a >= len) continue;
            const int c = std::min(len, a + r->seg_len) - a;
            cmax = std::max(cmax, c);
            R.pp += (int64_t)c * 64;
          }
          if (cmax > 0) work.push_back(WorkItemCost{BatchItem{(int32_t)k, (int32_t)t0, a, a + r->seg_len, j}, cmax});
        }
      } else {
        const int p_end = r->off + span;
        int cmax = 0;
        for (int64_t t = t0; t < t1; ++t) {
          const int len = r->buf.tile_len[(size_t)t];
          if (len <= r->off) continue;  // finished (or empty: its zeros are written at admission)
          const int c = std::min(len, p_end) - r->off;
          cmax = std::max(cmax, c);
          R.pp += (int64_t)c * 64;
        }
        if (cmax > 0) work.push_back(WorkItemCost{BatchItem{(int32_t)k, (int32_t)t0, r->off, p_end, -1}, cmax});
      }
    }
    // the request's progress once this round has run; the next round is built from it while this
    // one is still on the device
    r->off = r->nseg > 1 ? r->max_len : r->off + span;
    if (r->off >= r->max_len) {
      r->last_round = true;
      R.fin.push_back(r);
    }
  }
  build_queues(work, items, seg);
  return true;
}

void sarbp_server::launch_round(Round &R) {
  const int chunk = cfg.pulse_chunk;
  hipError_t e = hipSuccess;
  if (up_pending) {  // the requests admitted since the last round: their uploads first
    e = hipStreamWaitEvent(st, ev_up, 0);
    up_pending = false;
  }
  int rc = e == hipSuccess ? R.ib.upload(devs, items, seg, st, &segmented) : SARBP_OK;
  R.timed = false;
  if (!rc && e == hipSuccess && !items.empty()) {
    e = hipEventRecord(R.e0, st);
    // segmented requests (each in its only round): slots cleared first, means from their partial
    // sums after the batch kernel; one launch each for all of them (a kernel on st_up would hold
    // up the uploads queued behind it)
    if (e == hipSuccess) rc = sarbp_launch_segments_clear(R.ib.d_reqs, R.ib.d_list, (int)segmented.size(), st);
    if (!rc && e == hipSuccess)
      rc = sarbp_launch_batch(device, R.ib.d_reqs, R.ib.d_items, R.ib.d_seg, R.ib.d_heads, (int)items.size(),
                              sarbp_batch_mode(R.any_f64, R.any_split), R.waves, chunk, R.max_rows, 0, st);
    R.timed = !rc && e == hipSuccess;
  }
  if (!rc && e == hipSuccess) e = hipEventRecord(R.e1, st);  // after the round's batch kernel
  if (!rc && e == hipSuccess) e = hipEventRecord(R.ek, st);
  if (!rc && e == hipSuccess) e = hipStreamWaitEvent(st_comb, R.ek, 0);
  if (!rc && e == hipSuccess && !items.empty())
    rc = sarbp_launch_segments_combine(R.ib.d_reqs, R.ib.d_list, (int)segmented.size(), st_comb);
  if (!rc && e == hipSuccess) e = hipEventRecord(R.done, st_comb);
  R.failed = rc || e != hipSuccess;
  R.live = true;
}

void sarbp_server::fail_rounds(std::unique_lock<std::mutex> &lk) {
  lk.unlock();
  sync_all();  // nothing may still run when the rounds' blocks return to the pool
  lk.lock();
  for (Round &R : rounds) {
    if (!R.live) continue;
    for (ServerReq *r : R.reqs) {
      if (r->state != SARBP_REQ_RUNNING || !r->on_device) continue;  // already failed with the other round
      release(r);
      finish(r, SARBP_REQ_FAILED);
      stats.failed++;
    }
    R.live = false;
  }
  active.erase(std::remove_if(active.begin(), active.end(),
                              [](ServerReq *r) { return r->state != SARBP_REQ_RUNNING; }),
               active.end());
  cv_done.notify_all();
}

void sarbp_server::collect_round(Round &R, std::unique_lock<std::mutex> &lk) {
  if (!R.live) return;
  if (R.failed) {
    fail_rounds(lk);
    return;
  }
  lk.unlock();
  hipError_t e = hipEventSynchronize(R.done);
  float ms = 0.f;
  if (e == hipSuccess && R.timed) e = hipEventElapsedTime(&ms, R.e0, R.e1);
  lk.lock();
  if (e != hipSuccess) {
    fail_rounds(lk);
    return;
  }
  const double t_gpu = now();
  stats.rounds++;
  stats.gpu_busy_s += ms * 1e-3;
  stats.round_wall_s += t_gpu - R.t_round;
  if (ms > 0.05f && R.pp > 0) rate = 0.7 * rate + 0.3 * (R.pp * 0.8 / (ms * 1e-3));  // lanes ~80 % active
  for (ServerReq *r : R.reqs) {
    r->gpu_s = t_gpu;
    r->rounds++;
  }
  for (ServerReq *r : R.fin) {  // every tile wrote its means, now in the request's page-locked block
    release(r);
    r->completing = true;
    completing.push_back(r);
  }
  R.live = false;
  active.erase(std::remove_if(active.begin(), active.end(), [](ServerReq *r) { return r->completing; }),
               active.end());
  cv_complete.notify_all();
  cv_done.notify_all();
}

void sarbp_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  int next = 0;  // the round slot to build; the other one holds the round launched last
  bool admitted = false;
  for (;;) {
    // admission, FIFO: a request waits until a slot and its memory are free
    while (!stop && !queue.empty() && (int)active.size() < cfg.max_active) {
      ServerReq *r = queue.front();
      if (bytes_in_use + r->need > cfg.hbm_budget_bytes && !active.empty()) break;
      queue.pop_front();
      bytes_in_use += r->need;
      r->state = SARBP_REQ_RUNNING;
      lk.unlock();
      const double ta = now();
      const bool ok = admit(r);
      const double tb = now();
      lk.lock();
      stats.admit_s += tb - ta;
      if (!ok) {
        bytes_in_use -= r->need;
        finish(r, SARBP_REQ_FAILED);
        stats.failed++;
        cv_done.notify_all();
        continue;
      }
      if (r->admit_s < 0) r->admit_s = now();
      r->last_round = false;
      active.push_back(r);
      admitted = true;
    }
    if (admitted) {  // the next round's kernels wait for these uploads (only the uploads)
      admitted = false;
      if (hipEventRecord(ev_up, st_up) == hipSuccess) {
        up_pending = true;
      } else {
        (void)hipStreamSynchronize(st_up);
      }
    }
    if (stop) break;
    // requests asked to stop: their device contents are saved once no round that reads them is on
    // the device (a request in its last round completes instead)
    if (std::any_of(active.begin(), active.end(), [](ServerReq *r) { return r->preempt && !r->last_round; }))
      collect_round(rounds[next ^ 1], lk);
    for (auto it = active.begin(); it != active.end();) {
      ServerReq *r = *it;
      if (!r->preempt || r->last_round) {
        ++it;
        continue;
      }
      const size_t n = (size_t)r->buf.n;
      r->saved.resize(n);
      lk.unlock();
      hipError_t e = hipSuccess;
      if (up_pending) {  // a request admitted since the last round: its uploads first
        e = hipStreamWaitEvent(st, ev_up, 0);
        up_pending = false;
      }
      if (n && e == hipSuccess) {
        e = hipMemcpyAsync(r->saved.data(), r->buf.state, n * sizeof(BpState), hipMemcpyDeviceToHost, st);
      }
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      lk.lock();
      release(r);
      r->preempt = false;
      if (e == hipSuccess) {
        r->state = SARBP_REQ_PREEMPTED;
      } else {
        finish(r, SARBP_REQ_FAILED);
        stats.failed++;
      }
      r->preemptions++;
      it = active.erase(it);
      cv_done.notify_all();
    }
    Round &R = rounds[next];
    Round &P = rounds[next ^ 1];
    const double tb = now();
    if (!build_round(R)) {
      if (P.live) {
        collect_round(P, lk);
      } else {
        cv_work.wait_for(lk, std::chrono::milliseconds(20));
      }
      continue;
    }
    lk.unlock();
    launch_round(R);
    const double tl = now();
    lk.lock();
    stats.build_s += tl - tb;
    if (R.failed) {
      fail_rounds(lk);  // this round and the one before it
      continue;
    }
    collect_round(P, lk);  // wait for the previous round while this one is queued behind it
    next ^= 1;
  }
  // shutdown: the round on the device completes; what is left fails
  collect_round(rounds[next ^ 1], lk);
  for (ServerReq *r : active) {
    release(r);
    finish(r, SARBP_REQ_FAILED);
  }
  active.clear();
  for (ServerReq *r : queue) finish(r, SARBP_REQ_FAILED);
  queue.clear();
  cv_done.notify_all();
}

extern "C" {

int sarbp_server_create(const sarbp_server_config *cfg, sarbp_server **out) {
  if (!out) return sarbp_fail(SARBP_EINVAL, "out is null");
  *out = nullptr;
  sarbp_server *s = new (std::nothrow) sarbp_server();
  if (!s) return sarbp_fail(SARBP_ENOMEM, "host allocation failed");
  sarbp_server_config c{};
  if (cfg) c = *cfg;
  if (c.hbm_budget_bytes <= 0) c.hbm_budget_bytes = (int64_t)200e9;
  if (c.quantum_ms <= 0) c.quantum_ms = 20.0;
  if (c.max_active <= 0) c.max_active = 256;
  if (c.pulse_chunk <= 0) c.pulse_chunk = 32;
  if (c.phase_err_limit_rad <= 0) c.phase_err_limit_rad = 1e-3;
  if (c.max_segments <= 0) c.max_segments = 16;
  if (c.pulse_chunk % 8 || c.pulse_chunk > 1024) {
    delete s;
    return sarbp_fail(SARBP_EINVAL, "pulse_chunk must be a multiple of 8, at most 1024");
  }
  {
    const int rows = 64 / default_tiling(SARBP_PREC_SPLIT).tile_x;  // rows of the default split tile
    if (sarbp_batch_lds_bytes(1, c.pulse_chunk, rows) > SARBP_BATCH_LDS_MAX) {
      delete s;
      return sarbp_fail(SARBP_EINVAL, "pulse_chunk too large: its records for the default split tile's rows "
                                      "exceed the batch kernel's 64 KiB of LDS");
    }
  }
  s->cfg = c;
  hipError_t e = hipGetDevice(&s->device);
  int ncu = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->device);
  if (e == hipSuccess && ncu > 0) {
    s->slots_f64 = s->slots_split = (int64_t)ncu * 16;
    int b = 0;
    if (sarbp_batch_resident_blocks(s->device, SARBP_BATCH_F64, 1, c.pulse_chunk, 1, &b) == SARBP_OK && b > 0)
      s->slots_f64 = b;
    if (sarbp_batch_resident_blocks(s->device, SARBP_BATCH_SPLIT, 1, c.pulse_chunk, 4, &b) == SARBP_OK && b > 0)
      s->slots_split = b;
  }
  if (e == hipSuccess) e = hipStreamCreate(&s->st);
  if (e == hipSuccess) e = hipStreamCreate(&s->st_up);
  if (e == hipSuccess) e = hipStreamCreate(&s->st_comb);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_up, hipEventDisableTiming);
  for (Round &R : s->rounds) {
    if (e == hipSuccess) e = hipEventCreate(&R.e0);
    if (e == hipSuccess) e = hipEventCreate(&R.e1);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&R.ek, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&R.done, hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    int rc = sarbp_hip_fail(e, "server stream");
    s->destroy_events();
    s->destroy_streams();
    delete s;
    return rc;
  }
  // request blocks stay cached up to the budget or 16 GiB: freeing one waits for the device
  s->pool.limit = (size_t)std::min<int64_t>(c.hbm_budget_bytes, (int64_t)16 << 30);
  int started_done = 0;
  try {
    for (; started_done < sarbp_server::COMPLETION_THREADS; ++started_done)
      s->th_done[started_done] = std::thread([s] { s->complete_loop(); });
    s->th = std::thread([s] { s->run(); });
  } catch (...) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop_complete = true;
      s->cv_complete.notify_all();
    }
    for (int k = 0; k < started_done; ++k) s->th_done[k].join();
    s->destroy_events();
    s->destroy_streams();
    delete s;
    return sarbp_fail(SARBP_ENOMEM, "could not start the server's threads");
  }
  s->started = true;
  *out = s;
  return SARBP_OK;
}

int sarbp_server_submit(sarbp_server *s, sarbp_dataset *ds, const sarbp_request *req, double *out_c128,
                        int32_t *ncontrib, int64_t *id) {
  if (!s || !ds || !req || !id) return sarbp_fail(SARBP_EINVAL, "null argument");
  const double t_entry = s->now();  // a request's latency includes its submission
  if (req->n < 0 || (req->n > 0 && (!req->x || !req->y || !req->z || !req->mask || !req->index || !req->sint)))
    return sarbp_fail(SARBP_EINVAL, "bad request pixel arrays");
  if (req->n > 0 && !out_c128) return sarbp_fail(SARBP_EINVAL, "output buffer is null");
  if (req->dimx < 0 || (req->n > 0 && req->dimx > 0 && req->n % req->dimx))
    return sarbp_fail(SARBP_EINVAL, "request pixel count must be a multiple of dimx");
  if (req->precision != SARBP_PREC_F64 && req->precision != SARBP_PREC_SPLIT)
    return sarbp_fail(SARBP_EINVAL, "unknown request precision");
  const int tx = req->tile_x ? req->tile_x : default_tiling(req->precision).tile_x;
  if (tx != 64 && tx != 32 && tx != 16 && tx != 8 && tx != 4) return sarbp_fail(SARBP_EINVAL, "bad tile_x");
  auto r = std::make_unique<ServerReq>();
  r->ds = ds;
  r->out = out_c128;
  r->ncnt = ncontrib;
  r->req = *req;  // the scheduling decisions below read the caller's arrays
  r->precision = req->precision;
  r->req.tile_x = tx;
  // the request's page-locked block: its pixels in the device block's layout (uploaded at
  // admission with one copy, and again after a preemption), then the means and counts, which the
  // kernels write over PCIe as tiles finish (zero here: pixels no tile writes read as zero); its
  // layout does not depend on the state slots, which only the device block holds
  const int64_t nn = req->n;
  const size_t n = (size_t)nn;
  size_t off[10];
  reqbuf_layout(nn, 1, off);
  if (n) {
    r->out_off = off[6];
    r->host = s->hpool.get(off[8], &r->host_cap);
    if (!r->host) return sarbp_fail(SARBP_ENOMEM, "page-locked host allocation failed");
  }
  auto refuse = [&](int code, const char *msg) {
    s->hpool.put(r->host, r->host_cap);
    r->host = nullptr;
    return sarbp_fail(code, msg);
  };
  int split_rows = 0;  // rows per tile of a request that runs split (its LDS records per pulse)
  {
    // one pass over the pixels: tile unions, pixel-pulse pairs, the split rows' geometry, the
    // bounding boxes and the copy into the page-locked block, in bands of tile rows; a large request
    // splits the bands over threads of its own (about 2^16 pixels each)
    const bool split = req->precision == SARBP_PREC_SPLIT;
    const TileGrid tg = request_tile_grid(r->req);
    r->tile_len.assign((size_t)tg.ntiles(), 0);
    const int parts = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)SUBMIT_MAX_THREADS, nn >> 16, tg.nty}));
    std::vector<int64_t> pps((size_t)parts, 0);
    std::vector<SplitRows> rws((size_t)parts);
    std::vector<PixelBoxes> boxes((size_t)parts);
    char *h = r->host;
    auto band = [&](int k) {
      const int64_t by0 = tg.nty * k / parts, by1 = tg.nty * (k + 1) / parts;
      if (!n) return;
      request_tiles_band(ds, r->req, tg, by0, by1, r->tile_len.data(), pps[(size_t)k], split ? &rws[(size_t)k] : nullptr);
      const size_t i0 = (size_t)std::min<int64_t>(nn, by0 * tg.ty * tg.dimx);
      const size_t i1 = (size_t)std::min<int64_t>(nn, by1 * tg.ty * tg.dimx);
      boxes[(size_t)k].add(req->x, req->y, req->z, req->mask, req->sint, (int64_t)i0, (int64_t)i1);
      const size_t m = i1 - i0;
      std::memcpy(h + off[0] + i0 * 8, req->x + i0, m * 8);
      std::memcpy(h + off[1] + i0 * 8, req->y + i0, m * 8);
      std::memcpy(h + off[2] + i0 * 8, req->z + i0, m * 8);
      std::memcpy(h + off[3] + i0, req->mask + i0, m);
      std::memcpy(h + off[4] + i0 * 4, req->index + i0, m * 4);
      std::memcpy(h + off[5] + i0 * 4, req->sint + i0, m * 4);
      std::memset(h + off[6] + i0 * 16, 0, m * 16);
      std::memset(h + off[7] + i0 * 4, 0, m * 4);
      if (k == 0) {  // the alignment gaps after the means and the counts
        std::memset(h + off[6] + n * 16, 0, off[7] - off[6] - n * 16);
        std::memset(h + off[7] + n * 4, 0, off[8] - off[7] - n * 4);
      }
    };
    {
      std::vector<std::thread> th;
      int k = 1;
      try {
        for (; k < parts; ++k) th.emplace_back(band, k);
      } catch (...) {  // no thread to be had: the rest of the bands in this one
      }
      for (int j = k; j < parts; ++j) band(j);
      band(0);
      for (std::thread &t : th) t.join();
    }
    r->pp = 0;
    SplitRows rows;
    PixelBoxes box;
    for (int k = 0; k < parts; ++k) {
      r->pp += pps[(size_t)k];
      const SplitRows &q = rws[(size_t)k];
      rows.any = rows.any || q.any;
      rows.amax = std::max(rows.amax, q.amax);
      rows.omax = std::max(rows.omax, q.omax);
      box.merge(boxes[(size_t)k]);
    }
    double rmin = 0.0, rmax = 0.0;
    r->narrow = nn <= 0 || (sarbp_range_bounds_of(ds, box, rmin, rmax) && sarbp_narrow_ranges(ds, rmin, rmax));
    if (split) {
      rows.r0min = nn > 0 ? sarbp_r0min_of(ds, box) : 0.0;
      r->est = split_phase_estimate(ds, rows);
      r->r0min = (float)rows.r0min;
      if (r->est > s->cfg.phase_err_limit_rad) {  // FP64 terms, on the FP64 tiling
        r->precision = SARBP_PREC_F64;
        r->req.precision = SARBP_PREC_F64;
        r->req.tile_x = 0;
        TileGrid tg64;
        request_tiles(ds, r->req, tg64, r->tile_len, r->pp);
      } else {
        split_rows = tg.ty;
      }
    }
  }
  // a split request's records of one pulse chunk must fit the batch kernel's LDS; checked here so
  // that the request is refused alone instead of failing the round it would join
  if (split_rows > 0 && sarbp_batch_lds_bytes(1, s->cfg.pulse_chunk, split_rows) > SARBP_BATCH_LDS_MAX)
    return refuse(SARBP_EINVAL, "split request: pulse_chunk x tile rows x 48 B exceeds the batch kernel's "
                                "64 KiB of LDS (use a wider tile_x or a smaller pulse_chunk)");
  {
    int max_len = 0;
    int64_t live = 0;
    for (int v : r->tile_len) {
      max_len = std::max(max_len, v);
      live += v > 0;
    }
    const int64_t slots = r->precision == SARBP_PREC_SPLIT ? s->slots_split : s->slots_f64;
    r->nseg = segments_for(live, max_len, slots, s->cfg.pulse_chunk, s->cfg.max_segments, &r->seg_len);
  }
  // device bytes while admitted: pixels + outputs + state or segment slots (each array 256-byte
  // aligned)
  auto up = [](int64_t v) { return (v + 255) / 256 * 256; };
  r->need = up(nn * 8) * 3 + up(nn) + up(nn * 4) * 2 + up(nn * 16) + up(nn * 4) +
            up(nn * (int64_t)sizeof(BpState) * r->nseg);
  if (r->need > s->cfg.hbm_budget_bytes) return refuse(SARBP_EBUDGET, "request exceeds the device-memory budget");
  if (n) {
    char *h = r->host;
    r->req.x = (const double *)(h + off[0]);
    r->req.y = (const double *)(h + off[1]);
    r->req.z = (const double *)(h + off[2]);
    r->req.mask = (const uint8_t *)(h + off[3]);
    r->req.index = (const int32_t *)(h + off[4]);
    r->req.sint = (const int32_t *)(h + off[5]);
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) {
    s->hpool.put(r->host, r->host_cap);
    return sarbp_fail(SARBP_ESTATE, "server is shutting down");
  }
  r->id = s->next_id++;
  r->submit_s = t_entry;
  r->holds_ds = true;
  ds->server_refs.fetch_add(1);
  *id = r->id;
  s->queue.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  s->stats.submitted++;
  s->cv_work.notify_all();
  return SARBP_OK;
}

int sarbp_split_error_estimate(sarbp_dataset *ds, const sarbp_request *req, double *rms_phase_rad) {
  if (!ds || !req || !rms_phase_rad) return sarbp_fail(SARBP_EINVAL, "null argument");
  if (req->n < 0 || (req->n > 0 && (!req->x || !req->y || !req->z || !req->mask || !req->index || !req->sint)))
    return sarbp_fail(SARBP_EINVAL, "bad request pixel arrays");
  if (req->dimx < 0 || (req->n > 0 && req->dimx > 0 && req->n % req->dimx))
    return sarbp_fail(SARBP_EINVAL, "request pixel count must be a multiple of dimx");
  sarbp_request q = *req;
  q.tile_x = req->precision == SARBP_PREC_SPLIT ? req->tile_x : 0;
  q.precision = SARBP_PREC_SPLIT;
  const int tx = q.tile_x ? q.tile_x : default_tiling(SARBP_PREC_SPLIT).tile_x;
  if (tx != 64 && tx != 32 && tx != 16 && tx != 8 && tx != 4) return sarbp_fail(SARBP_EINVAL, "bad tile_x");
  TileGrid tg;
  std::vector<int> tile_len;
  int64_t pp = 0;
  SplitRows rows;
  request_tiles(ds, q, tg, tile_len, pp, &rows);
  *rms_phase_rad = split_phase_estimate(ds, rows);
  return SARBP_OK;
}

int sarbp_server_poll(sarbp_server *s, int64_t id, int32_t *state) {
  if (!s || !state) return sarbp_fail(SARBP_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  *state = it->second->state;
  return SARBP_OK;
}

int sarbp_server_wait(sarbp_server *s, int64_t id, double timeout_s) {
  if (!s) return sarbp_fail(SARBP_EINVAL, "null handle");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  ServerReq *r = it->second.get();
  auto fin = [r] { return r->state == SARBP_REQ_DONE || r->state == SARBP_REQ_FAILED || r->state == SARBP_REQ_PREEMPTED; };
  r->waiters++;  // sarbp_server_release leaves the record alone meanwhile
  bool ok = true;
  if (timeout_s < 0) {
    s->cv_done.wait(lk, fin);
  } else {
    ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), fin);
  }
  r->waiters--;
  if (!ok) return SARBP_EAGAIN;
  if (r->state == SARBP_REQ_FAILED) return sarbp_fail(SARBP_EDEVICE, "request failed");
  if (r->state == SARBP_REQ_PREEMPTED) return sarbp_fail(SARBP_ESTATE, "request is preempted");
  return SARBP_OK;
}

int sarbp_server_preempt(sarbp_server *s, int64_t id) {
  if (!s) return sarbp_fail(SARBP_EINVAL, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  ServerReq *r = it->second.get();
  if (r->state == SARBP_REQ_QUEUED) {
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), r), s->queue.end());
    r->state = SARBP_REQ_PREEMPTED;
    r->preemptions++;
    s->cv_done.notify_all();
    return SARBP_OK;
  }
  if (r->state != SARBP_REQ_RUNNING) return sarbp_fail(SARBP_ESTATE, "request is not queued or running");
  if (r->completing) return sarbp_fail(SARBP_ESTATE, "request is finished; its results are being delivered");
  r->preempt = true;
  s->cv_work.notify_all();
  return SARBP_OK;
}

int sarbp_server_resume(sarbp_server *s, int64_t id) {
  if (!s) return sarbp_fail(SARBP_EINVAL, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  ServerReq *r = it->second.get();
  if (r->state != SARBP_REQ_PREEMPTED) return sarbp_fail(SARBP_ESTATE, "request is not preempted");
  r->state = SARBP_REQ_QUEUED;
  s->queue.push_back(r);
  s->cv_work.notify_all();
  return SARBP_OK;
}

int sarbp_server_info(sarbp_server *s, int64_t id, sarbp_request_info *info) {
  if (!s || !info) return sarbp_fail(SARBP_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  const ServerReq *r = it->second.get();
  info->state = r->state;
  info->precision = r->precision;
  info->phase_err_est_rad = r->est;
  info->submit_s = r->submit_s;
  info->admit_s = r->admit_s;
  info->done_s = r->done_s;
  info->pp = r->pp;
  info->rounds = r->rounds;
  info->preemptions = r->preemptions;
  info->segments = r->nseg;
  info->round_s = r->round_s;
  info->gpu_done_s = r->gpu_s;
  return SARBP_OK;
}

int sarbp_server_release(sarbp_server *s, int64_t id) {
  if (!s) return sarbp_fail(SARBP_EINVAL, "null handle");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sarbp_fail(SARBP_EINVAL, "unknown request id");
  const ServerReq *r = it->second.get();
  if ((r->state != SARBP_REQ_DONE && r->state != SARBP_REQ_FAILED) || r->completing)
    return sarbp_fail(SARBP_ESTATE, "request is not finished");
  if (r->waiters > 0) return sarbp_fail(SARBP_ESTATE, "request is being waited on");
  s->reqs.erase(it);  // done or failed: no queue, round or completion list refers to it any more
  return SARBP_OK;
}

int sarbp_server_get_stats(sarbp_server *s, sarbp_server_stats *st) {
  if (!s || !st) return sarbp_fail(SARBP_EINVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  *st = s->stats;
  st->device_bytes_in_use = s->bytes_in_use;
  st->active = (int32_t)s->active.size();
  st->queued = (int32_t)s->queue.size();
  st->records = (int64_t)s->reqs.size();
  return SARBP_OK;
}

int sarbp_server_destroy(sarbp_server *s) {
  if (!s) return SARBP_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->started && s->th.joinable()) s->th.join();
  {  // the completion threads finish the copies still queued, then stop
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop_complete = true;
    s->cv_complete.notify_all();
  }
  for (auto &t : s->th_done)
    if (t.joinable()) t.join();
  s->sync_all();
  for (Round &R : s->rounds) R.ib.release();
  s->pool.release();
  for (auto &kv : s->reqs) {  // preempted requests never resumed still hold their host copy
    if (kv.second->host) (void)hipHostFree(kv.second->host);
    sarbp_server::drop_ds(kv.second.get());
  }
  s->hpool.release();
  s->destroy_events();
  s->destroy_streams();
  delete s;
  return SARBP_OK;
}

}  // extern "C"
