// server.hip -- continuous-batching request loop (hbj_server_*).
//
// Host threads share one device (an admitter and one worker per lane):
//   admitter  takes queued requests in arrival order.  For each it first
//             builds the system's host image (tile tables, the sliced SpMV
//             layout, the LU blocks, on the host), which fixes the exact
//             size of its device block; then, once the HBM budget and the
//             slot count allow, places the block in the server's device
//             arena and uploads it: the image goes into pinned staging
//             (two buffers, alternating), one asynchronous copy moves it on
//             the server's upload stream and an event marks its end.  A
//             system with the sliced layout keeps only that layout on the
//             device (no CSR arrays, no per-entry indices).  This host work
//             and the copies overlap the iterations running on the GPU;
//   workers   one per lane (two lanes: two tables on two streams, side by
//             side, so that each table's launches fill the other's launch
//             tails and gaps; the admitter puts each request in the lane
//             with fewer resident tiles).  A worker owns its lane's slot
//             table.  It enqueues chunks of PCG iterations
//             (one launch of each kernel covers every resident request) and
//             always has the next chunk queued before it waits for the
//             previous one's status poll.  A new request joins the table
//             once its upload event has completed (polled between chunks;
//             waited on only when the lane has nothing else to run), so
//             the resident requests' iterations never queue behind an upload.
//             The SpMV reads the values through the caches (variant 5) while
//             the resident systems' layouts fit the last-level cache, and
//             non-temporally (variant 4) otherwise.  Arrivals and retirements change
//             the table between two enqueues, without waiting for the
//             device: the slot and tile writes are copies on the engine's
//             stream, so they land between the chunk in flight and the next.
//             A finished request's x is copied out at once; its system's
//             block goes back to the arena once the last chunk that may
//             still reference the system has been polled (no hipFree, which
//             would wait for the device).
// Each request's arithmetic only depends on its own system (tile partials and
// last-tile reductions are per system), so its result is the same alone or
// in any mix.
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "engine.hpp"

namespace {

struct Request {
  uint64_t ticket = 0;
  uint32_t n = 0, nnz = 0, n_starts = 0;
  std::vector<uint32_t> rp, col, bs;
  std::vector<float> val;
  std::vector<double> b64;
  hbj_ir_params params{};
  uint64_t est_bytes = 0;      // the exact device block, once the image is built
  double t_submit = 0, t_done = 0;
  hbj_system* sys = nullptr;
  hipEvent_t up_ev = nullptr;  // the end of the system's upload
  int slot = -1;
  int lane = 0;
  bool done = false;
  int32_t status = HBJ_OK;
  std::string error;
  hbj_result res{};
  std::vector<double> x;
};

}  // namespace

struct hbj_server {
  hbj_ctx* ctx = nullptr;
  hbj_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_admit, cv_work, cv_done;
  std::deque<Request*> queue;     // submitted, not yet admitted
  static constexpr int kLanes = 2;
  int n_lanes = 1;
  std::deque<Request*> ready[kLanes];  // resident on the device, not yet in a lane's table
  uint64_t lane_tiles[kLanes] = {0, 0};   // tiles of each lane's resident requests
  uint32_t lane_active[kLanes] = {0, 0};  // requests in each lane's table
  std::unordered_map<uint64_t, Request*> reqs;
  uint64_t next_ticket = 1;
  uint64_t bytes_in_use = 0;      // reserved by admitted requests
  uint32_t resident = 0;          // admitted and not yet released
  DeviceArena arena;              // device memory of the resident systems
  uint64_t arena_gives = 0;       // systems whose blocks went back to the arena
  bool arena_blocked = false;     // the queue head found no free range ...
  uint64_t blocked_at = 0;        // ... when arena_gives had this value
  hbj_server_stats st{};
  bool stop = false;
  bool failed = false;            // the worker stopped on a device failure
  std::thread admitter, worker[kLanes];
  Engine eng[kLanes];
  // uploads: the admitter's stream, pinned staging buffers (each free once
  // its event has completed) and the buffers they outgrew (freed at the end:
  // hipHostFree waits for all device work)
  hipStream_t up_stream = nullptr;
  char* stg[2] = {nullptr, nullptr};
  uint64_t stg_cap[2] = {0, 0};
  hipEvent_t stg_ev[2] = {nullptr, nullptr};
  bool stg_used[2] = {false, false};
  int stg_next = 0;
  std::vector<char*> stg_old;
};

// The SpMV reads the values through the caches (variant 5) while the
// resident systems' sliced layouts total at most this many bytes (the
// last-level cache holds 256 MB; the vectors the iterations stream need the
// rest), non-temporally (variant 4) above it.
constexpr uint64_t kCachedLayoutBytes = 160ull << 20;

namespace {

void finish(hbj_server* S, Request* r, int32_t status, const std::string& err) {
  r->status = status;
  r->error = err;
  r->t_done = now_ms();
  r->done = true;
  S->st.completed += 1;
  S->cv_done.notify_all();
}

// A pinned staging buffer of at least `bytes` that no upload still reads.
hbj_status staging_take(hbj_server* S, uint64_t bytes, char** out, int* slot) {
  const int k = S->stg_next;
  S->stg_next ^= 1;
  if (S->stg_used[k]) HIPCHK(hipEventSynchronize(S->stg_ev[k]));
  if (S->stg_cap[k] < bytes) {
    if (S->stg[k]) S->stg_old.push_back(S->stg[k]);
    S->stg[k] = nullptr;
    S->stg_cap[k] = 0;
    const uint64_t cap = bytes + bytes / 4 + (1ull << 20);
    CHK(halloc(&S->stg[k], cap));
    S->stg_cap[k] = cap;
  }
  *out = S->stg[k];
  *slot = k;
  return HBJ_OK;
}

void admitter_main(hbj_server* S) {
  (void)hipSetDevice(S->ctx->device);
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    S->cv_admit.wait(lk, [&] { return S->failed || S->stop || !S->queue.empty(); });
    if (S->failed || S->queue.empty()) break;  // device failure, or stop requested and nothing left
    Request* r = S->queue.front();
    S->queue.pop_front();
    lk.unlock();
    // the host image: tables, sliced layout, LU blocks (no device memory)
    const double t0 = now_ms();
    SystemImage im;
    hbj_status s = prepare_image(r->n, r->nnz, r->rp.data(), r->col.data(), r->val.data(),
                                 r->bs.data(), r->n_starts, true, im);
    std::string err = s == HBJ_OK ? "" : hbj_last_error();
    const uint64_t need = s == HBJ_OK ? lean_block_bytes(im) : 0u;
    lk.lock();
    S->st.admit_ms += now_ms() - t0;
    if (s != HBJ_OK) {
      finish(S, r, s, err);
      continue;
    }
    if (need > S->cfg.hbm_budget_bytes) {
      S->st.rejected += 1;
      char msg[200];
      snprintf(msg, sizeof msg, "request needs %llu device bytes, above the admission budget of %llu",
               (unsigned long long)need, (unsigned long long)S->cfg.hbm_budget_bytes);
      finish(S, r, HBJ_E_REJECTED, msg);
      continue;
    }
    r->est_bytes = need;
    // budget, a slot, and after a failed placement a block given back
    uint64_t off = 0;
    bool placed = false;
    while (!S->failed) {
      S->cv_admit.wait(lk, [&] {
        if (S->failed) return true;
        if (S->arena_blocked && S->arena_gives == S->blocked_at) return false;
        return S->resident < S->cfg.max_active && S->bytes_in_use + need <= S->cfg.hbm_budget_bytes;
      });
      if (S->failed) break;
      S->arena_blocked = false;
      if (S->arena.take(need, &off)) {
        placed = true;
        break;
      }
      // the arena's free space is split: wait until a resident system gives
      // its block back (there is one: an empty arena holds the whole budget)
      S->arena_blocked = true;
      S->blocked_at = S->arena_gives;
      S->st.arena_waits += 1;
    }
    if (!placed) {
      finish(S, r, HBJ_E_DEVICE, "server stopped on a device failure");
      break;
    }
    S->bytes_in_use += need;
    S->resident += 1;
    lk.unlock();
    // upload: the image into pinned staging, one asynchronous copy
    const double t1 = now_ms();
    hbj_system* sys = nullptr;
    hipEvent_t ev = nullptr;
    char* stg = nullptr;
    int slot = 0;
    uint64_t copied = 0;
    s = staging_take(S, need, &stg, &slot);
    if (s == HBJ_OK && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipEventCreate failed");
      s = HBJ_E_DEVICE;
    }
    if (s == HBJ_OK)
      s = system_place(S->ctx, im, r->rp.data(), r->col.data(), r->val.data(), r->bs.data(),
                       r->b64.data(), &S->arena, off, stg, S->up_stream, ev, &sys, &copied);
    if (s == HBJ_OK && hipEventRecord(S->stg_ev[slot], S->up_stream) != hipSuccess) {
      (void)hipGetLastError();
      set_error("hipEventRecord failed");
      s = HBJ_E_DEVICE;
    }
    if (s == HBJ_OK) S->stg_used[slot] = true;
    err = s == HBJ_OK ? "" : hbj_last_error();
    lk.lock();
    S->st.admit_ms += now_ms() - t1;
    if (s != HBJ_OK) {
      if (sys) {
        sys->arena = nullptr;
        delete sys;
      }
      if (ev) (void)hipEventDestroy(ev);
      S->arena.give(off, need);
      S->bytes_in_use -= need;
      S->resident -= 1;
      finish(S, r, s, err);
      S->cv_admit.notify_all();
      continue;
    }
    S->st.upload_bytes += copied;
    // host copies are no longer needed once the image is staged
    r->sys = sys;
    r->up_ev = ev;
    std::vector<uint32_t>().swap(r->rp);
    std::vector<uint32_t>().swap(r->col);
    std::vector<uint32_t>().swap(r->bs);
    std::vector<float>().swap(r->val);
    std::vector<double>().swap(r->b64);
    int lane = 0;
    for (int k = 1; k < S->n_lanes; ++k)
      if (S->lane_tiles[k] < S->lane_tiles[lane]) lane = k;
    r->lane = lane;
    S->lane_tiles[lane] += sys->n_tiles;
    S->ready[lane].push_back(r);
    if (S->bytes_in_use > S->st.peak_device_bytes) S->st.peak_device_bytes = S->bytes_in_use;
    S->cv_work.notify_all();
  }
}

// Worker: returns a status only on a device failure (all requests then fail).
hbj_status worker_loop(hbj_server* S, int lane) {
  HIPCHK(hipSetDevice(S->ctx->device));
  Engine& E = S->eng[lane];
  std::deque<Request*>& ready = S->ready[lane];
  // the lanes share the last-level cache
  const uint64_t cached_limit = kCachedLayoutBytes / (uint64_t)S->n_lanes;
  const uint32_t cap = S->cfg.max_active;
  const uint32_t T = S->cfg.chunk_iters ? S->cfg.chunk_iters : 64;
  // Chunks are numbered in enqueue order.  A request placed in a slot is
  // advanced from chunk since[slot] on, so polls of earlier chunks (which
  // show the slot's previous request) are not read for it.
  std::vector<Request*> slots(cap, nullptr);
  std::vector<uint64_t> since(cap, 0);
  // Retired systems: a chunk enqueued before the retirement may still
  // reference them; their blocks go back once chunk `after` has been polled.
  struct Grave {
    hbj_system* sys;
    uint64_t bytes;
    uint64_t after;
  };
  std::vector<Grave> graves;
  uint64_t next_chunk = 0;
  int inflight = -1;  // pool of the chunk whose poll is pending
  uint64_t inflight_id = 0;
  int cur = 0;
  uint32_t active = 0;
  bool table_dirty = false;
  uint64_t cached_chunks = 0;
  E.set_spmv_variant(4u);

  // retire every slot whose polled status is final
  auto harvest = [&](const SysState* polled, uint64_t chunk_id) -> hbj_status {
    for (uint32_t i = 0; i < cap; ++i) {
      Request* r = slots[i];
      if (!r || since[i] > chunk_id || polled[i].status == kRunning) continue;
      hbj_result res{};
      const SysState& s = polled[i];
      res.status = s.status;
      res.iterations = s.total_iter;
      res.outer_iterations = s.outer;
      res.rr = s.rr;
      res.b_norm2 = s.b_norm2;
      res.rel_res_recurrence = s.b_norm2 > 0 ? std::sqrt((double)s.rr / (double)s.b_norm2) : 0.0;
      res.rel_res_true_fp64 = s.bb64 > 0 ? std::sqrt(s.rr64 / s.bb64) : 0.0;
      r->x.assign(r->n, 0.0);
      // the chunks still queued no longer write this system (its status is final)
      CHK(hbj_system_get_x64(r->sys, r->x.data()));
      graves.push_back(Grave{r->sys, r->est_bytes, next_chunk - 1});
      if (r->up_ev) (void)hipEventDestroy(r->up_ev);  // long complete: the system ran
      r->up_ev = nullptr;
      std::lock_guard<std::mutex> g(S->mu);
      S->lane_tiles[lane] -= r->sys->n_tiles;
      r->res = res;
      r->res.solve_ms = now_ms() - r->t_submit;
      std::string err;
      if (s.status == HBJ_E_NOT_CONVERGED)
        err = "IR: relative residual above target after the refinement limit";
      else if (s.status == HBJ_E_BREAKDOWN_PAP)
        err = "PCG(BlockJacobiGpu): dot(p,Ap) is zero (breakdown)";
      else if (s.status == HBJ_E_BREAKDOWN_RZ)
        err = "PCG(BlockJacobiGpu): rz_old is zero (breakdown)";
      r->sys = nullptr;
      slots[i] = nullptr;
      finish(S, r, s.status, err);  // r is the client's from here on
      --active;
      table_dirty = true;
    }
    return HBJ_OK;
  };
  // return the blocks of systems no chunk after `polled_id` references
  auto bury = [&](uint64_t polled_id) {
    size_t keep = 0;
    uint64_t freed = 0, count = 0;
    for (const Grave& g : graves) {
      if (g.after <= polled_id) {
        hbj_system_destroy(g.sys);  // an arena system: no HIP call
        freed += g.bytes;
        ++count;
      } else {
        graves[keep++] = g;
      }
    }
    graves.resize(keep);
    if (!count) return;
    std::lock_guard<std::mutex> g(S->mu);
    S->bytes_in_use -= freed;
    S->resident -= (uint32_t)count;
    S->arena_gives += count;
    S->cv_admit.notify_all();
    S->cv_work.notify_all();  // the other lanes may be waiting to stop
  };

  // admitted requests whose upload is still in flight: they join the table
  // once it has landed, so the lane's iterations never wait for an upload
  std::vector<Request*> uploading;
  for (;;) {
    std::vector<Request*> incoming;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (active == 0 && inflight < 0 && uploading.empty()) {
        S->cv_work.wait(lk, [&] { return !ready.empty() || (S->stop && S->queue.empty() && S->resident == 0); });
        if (ready.empty() && S->stop && S->queue.empty() && S->resident == 0) break;
      }
      while (!ready.empty()) {
        uploading.push_back(ready.front());
        ready.pop_front();
      }
    }
    {
      size_t keep = 0;
      for (Request* r : uploading) {
        hipError_t q = hipEventQuery(r->up_ev);
        // nothing else to run: wait for the first upload rather than spin
        if (q == hipErrorNotReady && active == 0 && inflight < 0 && incoming.empty())
          q = hipEventSynchronize(r->up_ev);
        if (q == hipSuccess) incoming.push_back(r);
        else if (q == hipErrorNotReady) uploading[keep++] = r;
        else {
          set_error("upload event: %s", hipGetErrorString(q));
          return HBJ_E_DEVICE;
        }
      }
      uploading.resize(keep);
    }
    bool staged = false;
    if (!incoming.empty() || table_dirty) {
      const double t0 = now_ms();
      CHK(E.stage_begin(cur));
      for (Request* r : incoming) {
        uint32_t slot = 0;
        while (slots[slot]) ++slot;  // resident <= cap guarantees a free slot
        slots[slot] = r;
        since[slot] = next_chunk;
        r->slot = (int)slot;
        // the upload has landed (its event completed); the stream-ordered
        // wait only keeps the order explicit
        HIPCHK(hipStreamWaitEvent(E.stream, r->up_ev, 0));
        CHK(E.stage_slot(cur, slot, r->sys->desc(true, E.fused),
                         r->sys->init_state(true, 0, 0, 0, &r->params, false, E.fused)));
        ++active;
      }
      std::vector<uint2> tiles;
      uint64_t layouts = 0;
      for (uint32_t i = 0; i < cap; ++i)
        if (slots[i]) {
          for (uint32_t t = 0; t < slots[i]->sys->n_tiles; ++t) tiles.push_back(make_uint2(i, t));
          layouts += slots[i]->sys->spmv_sliced_bytes ? slots[i]->sys->spmv_sliced_bytes
                                                      : slots[i]->sys->spmv_layout_bytes;
        }
      CHK(E.stage_tiles(cur, tiles));
      // the resident layouts' bytes pick the SpMV's cache policy (variants 4
      // and 5 are the only ones a system without CSR arrays can run)
      E.set_spmv_variant(layouts <= cached_limit ? 5u : 4u);
      table_dirty = false;
      staged = true;
      std::lock_guard<std::mutex> g(S->mu);
      S->lane_active[lane] = active;
      S->st.active = 0;
      for (int k = 0; k < S->n_lanes; ++k) S->st.active += S->lane_active[k];
      if (S->st.active > S->st.peak_active) S->st.peak_active = S->st.active;
      S->st.table_changes += 1;
      S->st.table_ms += now_ms() - t0;
    }
    int enq = -1;
    uint64_t enq_id = 0;
    if (active > 0) {
      CHK(E.enqueue_poll(T, cur));
      enq = cur;
      enq_id = next_chunk++;
      if (E.spmv_variant == 5u) cached_chunks = 1;
    } else if (staged) {
      CHK(E.sync_stream());  // nothing follows the staged copies: let them land
    }
    if (inflight >= 0) {
      CHK(E.wait_poll(inflight));
      CHK(harvest(E.h_poll[inflight], inflight_id));
      bury(inflight_id);
      inflight = -1;
    }
    if (enq >= 0) {
      inflight = enq;
      inflight_id = enq_id;
      cur ^= 1;
    }
    std::lock_guard<std::mutex> g(S->mu);
    S->lane_active[lane] = active;
    S->st.active = 0;
    for (int k = 0; k < S->n_lanes; ++k) S->st.active += S->lane_active[k];
    if (enq >= 0) S->st.chunks += 1;
    S->st.chunks_cached_spmv += cached_chunks;
    cached_chunks = 0;
  }
  return HBJ_OK;
}

void worker_main(hbj_server* S, int lane) {
  hbj_status s = worker_loop(S, lane);
  if (s != HBJ_OK) {
    // device failure: fail everything that is still outstanding
    const std::string err = hbj_last_error();
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;
    S->failed = true;
    for (auto& kv : S->reqs)
      if (!kv.second->done) finish(S, kv.second, s, err);
    S->cv_admit.notify_all();
  }
}

// The upload stream, staging buffers and their events (after the stream's
// copies are done).
void release_uploads(hbj_server* S) {
  if (S->up_stream) (void)hipStreamSynchronize(S->up_stream);
  for (int k = 0; k < 2; ++k) {
    hfree(S->stg[k]);
    if (S->stg_ev[k]) (void)hipEventDestroy(S->stg_ev[k]);
    S->stg_ev[k] = nullptr;
  }
  for (char*& p : S->stg_old) hfree(p);
  S->stg_old.clear();
  if (S->up_stream) (void)hipStreamDestroy(S->up_stream);
  S->up_stream = nullptr;
}

// Device bytes no request of these sizes can go below (its nonzero values
// once -- every layout stores those, a sliced one may leave out the +-0
// entries --, its LU blocks and block starts, the right-hand side and the
// solver's vectors): requests above the budget by this much are rejected at
// once, the others once their exact block is known.
uint64_t min_request_bytes(uint32_t n, uint32_t nnz, const float* values, uint32_t n_starts) {
  const uint64_t nb = n_starts ? n_starts - 1u : 0u;
  uint64_t stored = 0;
  for (uint32_t e = 0; e < nnz; ++e) stored += values[e] != 0.0f ? 1u : 0u;
  return 4ull * stored + 76ull * n + 144ull * nb + 4ull * n_starts;
}

}  // namespace

extern "C" {

hbj_status hbj_server_create(hbj_ctx* ctx, const hbj_server_config* cfg, hbj_server** out) {
  clear_error();
  if (!ctx || !cfg || !out || cfg->max_active == 0 || cfg->hbm_budget_bytes == 0) {
    set_error("hbj_server_create: need a context, max_active > 0 and hbm_budget_bytes > 0");
    return HBJ_E_INVALID;
  }
  auto* S = new hbj_server();
  S->ctx = ctx;
  S->cfg = *cfg;
  if (S->cfg.chunk_iters == 0) S->cfg.chunk_iters = 64;
  // even: a system's state copy 0 is current at every chunk boundary
  S->cfg.chunk_iters = S->cfg.chunk_iters < 2 ? 2u : (S->cfg.chunk_iters & ~1u);
  hipError_t e = hipSetDevice(ctx->device);
  // tile table sized for the budget (a 256-row tile holds at least 72 bytes
  // per row of vectors), at most 2^18 tiles to start; it grows if needed
  const uint64_t tiles0 = std::min<uint64_t>(cfg->hbm_budget_bytes / (72ull * kTile) + 1, 1ull << 18);
  {
    const uint32_t lanes = env_u32("HBJ_SERVER_LANES", 2);
    S->n_lanes = lanes < 1 ? 1 : (lanes > (uint32_t)hbj_server::kLanes ? hbj_server::kLanes : (int)lanes);
  }
  hbj_status s = e == hipSuccess ? HBJ_OK : HBJ_E_DEVICE;
  for (int k = 0; k < S->n_lanes && s == HBJ_OK; ++k) {
    Engine& E = S->eng[k];
    s = E.init(ctx->device, cfg->max_active, (uint32_t)tiles0);
    if (s != HBJ_OK) break;
    E.use_graph = false;  // the table changes between chunks
    E.fused = fused_tables();  // a table of many systems: three launches
    E.n_tiles = 0;
    // slots start zeroed (status 0: final); no tile references a free slot
    const uint32_t cap = E.cap;
    e = hipMemset(E.d_descs, 0, sizeof(SysDesc) * cap);
    if (e == hipSuccess) e = hipMemset(E.d_states, 0, sizeof(SysState) * 2 * cap);
    if (e != hipSuccess) {
      set_error("hbj_server_create: %s", hipGetErrorString(e));
      s = HBJ_E_DEVICE;
    }
  }
  // the resident systems' device memory: one allocation of the budget
  if (s == HBJ_OK) s = S->arena.init(cfg->hbm_budget_bytes);
  // uploads: a stream of their own and the staging buffers' events
  if (s == HBJ_OK) {
    e = hipStreamCreateWithFlags(&S->up_stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      e = hipEventCreateWithFlags(&S->stg_ev[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error("hbj_server_create: %s", hipGetErrorString(e));
      s = HBJ_E_DEVICE;
    }
  }
  if (s != HBJ_OK) {
    release_uploads(S);
    for (int k = 0; k < hbj_server::kLanes; ++k) S->eng[k].destroy();
    S->arena.destroy();
    delete S;
    return s;
  }
  S->admitter = std::thread(admitter_main, S);
  for (int k = 0; k < S->n_lanes; ++k) S->worker[k] = std::thread(worker_main, S, k);
  *out = S;
  return HBJ_OK;
}

void hbj_server_destroy(hbj_server* S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->stop = true;  // outstanding requests still run to completion
  }
  S->cv_admit.notify_all();
  S->cv_work.notify_all();
  if (S->admitter.joinable()) S->admitter.join();
  {
    std::lock_guard<std::mutex> g(S->mu);
    S->cv_work.notify_all();
  }
  for (int k = 0; k < hbj_server::kLanes; ++k)
    if (S->worker[k].joinable()) S->worker[k].join();
  (void)hipSetDevice(S->ctx->device);
  for (auto& kv : S->reqs) {
    if (kv.second->sys) hbj_system_destroy(kv.second->sys);
    if (kv.second->up_ev) (void)hipEventDestroy(kv.second->up_ev);
    delete kv.second;
  }
  release_uploads(S);
  for (int k = 0; k < hbj_server::kLanes; ++k) S->eng[k].destroy();
  S->arena.destroy();
  delete S;
}

hbj_status hbj_server_submit(hbj_server* S, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                             const uint32_t* col_idx, const float* values,
                             const uint32_t* block_starts, uint32_t n_starts, const double* b64,
                             const hbj_ir_params* params, uint64_t* ticket) {
  clear_error();
  if (!S || !row_ptr || !block_starts || !params || !ticket || (n && !b64) ||
      (nnz && (!col_idx || !values))) {
    set_error("hbj_server_submit: bad argument");
    return HBJ_E_INVALID;
  }
  if (params->max_inner == 0 || !(params->inner_rel_tol > 0.0f) || !(params->target_rel_res > 0.0)) {
    set_error("hbj_ir_params: max_inner > 0, inner_rel_tol > 0 and target_rel_res > 0 required");
    return HBJ_E_INVALID;
  }
  CHK(validate_case(n, nnz, row_ptr, col_idx, block_starts, n_starts));
  auto* r = new Request();
  r->n = n;
  r->nnz = nnz;
  r->n_starts = n_starts;
  const uint64_t min_bytes = min_request_bytes(n, nnz, values, n_starts);
  std::lock_guard<std::mutex> g(S->mu);
  if (S->stop) {
    delete r;
    set_error("hbj_server_submit: server is shutting down");
    return HBJ_E_INVALID;
  }
  if (min_bytes > S->cfg.hbm_budget_bytes) {
    S->st.rejected += 1;
    set_error("request needs at least %llu device bytes, above the admission budget of %llu",
              (unsigned long long)min_bytes, (unsigned long long)S->cfg.hbm_budget_bytes);
    delete r;
    return HBJ_E_REJECTED;
  }
  r->rp.assign(row_ptr, row_ptr + n + 1);
  r->col.assign(col_idx, col_idx + nnz);
  r->val.assign(values, values + nnz);
  r->bs.assign(block_starts, block_starts + n_starts);
  r->b64.assign(b64, b64 + n);
  r->params = *params;
  r->ticket = S->next_ticket++;
  r->t_submit = now_ms();
  S->reqs[r->ticket] = r;
  S->queue.push_back(r);
  S->st.submitted += 1;
  *ticket = r->ticket;
  S->cv_admit.notify_all();
  return HBJ_OK;
}

hbj_status hbj_server_wait(hbj_server* S, uint64_t ticket, double* x, hbj_result* result,
                           double* latency_ms) {
  clear_error();
  if (!S) {
    set_error("hbj_server_wait: null server");
    return HBJ_E_INVALID;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) {
    set_error("hbj_server_wait: unknown ticket %llu", (unsigned long long)ticket);
    return HBJ_E_INVALID;
  }
  Request* r = it->second;
  S->cv_done.wait(lk, [&] { return r->done; });
  if (x && r->status == HBJ_OK) std::memcpy(x, r->x.data(), sizeof(double) * r->n);
  if (result) *result = r->res;
  if (result) result->status = r->status;
  if (latency_ms) *latency_ms = r->t_done - r->t_submit;
  const int32_t st = r->status;
  if (st != HBJ_OK) set_error("%s", r->error.c_str());
  S->reqs.erase(it);
  delete r;
  return (hbj_status)st;
}

uint64_t hbj_server_request_bytes(uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                                  const uint32_t* col_idx, const float* values,
                                  const uint32_t* block_starts, uint32_t n_starts) {
  clear_error();
  if (!row_ptr || !block_starts || (nnz && (!col_idx || !values))) {
    set_error("hbj_server_request_bytes: bad argument");
    return 0;
  }
  if (validate_case(n, nnz, row_ptr, col_idx, block_starts, n_starts) != HBJ_OK) return 0;
  SystemImage im;
  if (prepare_image(n, nnz, row_ptr, col_idx, values, block_starts, n_starts, true, im) != HBJ_OK)
    return 0;
  return lean_block_bytes(im);
}

hbj_status hbj_server_stats_get(hbj_server* S, hbj_server_stats* out) {
  if (!S || !out) {
    set_error("hbj_server_stats_get: bad argument");
    return HBJ_E_INVALID;
  }
  std::lock_guard<std::mutex> g(S->mu);
  *out = S->st;
  out->queued = S->queue.size();
  for (int k = 0; k < S->n_lanes; ++k) out->queued += S->ready[k].size();
  out->device_bytes_in_use = S->bytes_in_use;
  return HBJ_OK;
}

}  // extern "C"
