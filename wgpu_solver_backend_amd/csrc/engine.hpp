// engine.hpp -- host-side objects shared by runtime.hip and server.hip:
// allocation helpers, the context, the iteration engine and the
// device-resident system.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hbj_internal.hpp"
#include "ranges.hpp"

using namespace hbj;

#define HIPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      set_error("%s failed: %s", #expr, hipGetErrorString(e_));              \
      return HBJ_E_DEVICE;                                                    \
    }                                                                         \
  } while (0)

#define CHK(expr)                    \
  do {                               \
    hbj_status s_ = (expr);          \
    if (s_ != HBJ_OK) return s_;     \
  } while (0)

namespace hbj_detail {

template <typename T>
inline hbj_status dalloc(T** p, size_t count, uint64_t* acct = nullptr) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
    return HBJ_E_OOM;
  }
  if (acct) *acct += count * sizeof(T);
  return HBJ_OK;
}

template <typename T>
inline void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

template <typename T>
inline hbj_status halloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipHostMalloc((void**)p, count * sizeof(T), hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error("hipHostMalloc(%zu bytes) failed: %s", count * sizeof(T), hipGetErrorString(e));
    return HBJ_E_OOM;
  }
  return HBJ_OK;
}

template <typename T>
inline void hfree(T*& p) {
  if (p) (void)hipHostFree((void*)p);
  p = nullptr;
}

inline double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

inline uint32_t env_u32(const char* name, uint32_t dflt) {
  const char* v = std::getenv(name);
  if (!v || !*v) return dflt;
  return (uint32_t)std::strtoul(v, nullptr, 10);
}

// A rocprofiler-sdk tool (rocprofv3 and the like) is loaded into the process:
// the SDK loads the libraries named by ROCP_TOOL_LIBRARIES.  Replaying a
// captured graph under such a tool's kernel tracing crashed on the host
// inside hipGraphLaunch (a read of an unmapped address by the runtime's
// intercepted dispatch path; the same graphs replay correctly without the
// tool), so the engine launches the kernels directly while one is attached.
// The two-launch PCG iteration (k_pcg_fspmv + k_pcg_update) or the
// three-launch one (k_pcg_spmv, k_pcg_update, k_pcg_finish): same arithmetic,
// same bits.  By default a single system of at most kFetchMaxTiles tiles runs
// two launches (50^3 lattice: 32.9 against 35.2 us per iteration) and a table
// of several systems three (its finish kernel reduces each system's partials
// once per four tiles, and its SpMV holds fewer registers: 256-system batch
// 828 against 777 solves/s, serving 67.5 against 63.8 requests/s);
// HBJ_FUSED=1 selects two launches for tables too (where their systems
// allow it), HBJ_FUSED=0 three everywhere.  Read when an engine is set up (a
// later change applies to engines set up after it).
inline uint32_t fused_env() { return env_u32("HBJ_FUSED", 2); }
inline bool fused_mode() { return fused_env() != 0; }    // a single system
inline bool fused_tables() { return fused_env() == 1; }  // a table of several

inline bool profiler_tool_attached() {
  const char* v = std::getenv("ROCP_TOOL_LIBRARIES");
  return v && *v;
}

}  // namespace hbj_detail
using namespace hbj_detail;

// ------------------------------------------------------------------ context

struct hbj_ctx {
  int device = 0;
  hipDeviceProp_t prop{};
  hipStream_t stream = nullptr;  // single-op executors
  std::mutex mu;
};

// ------------------------------------------------------------------ engine

namespace hbj {

// Drives the three PCG kernels over a tile table until every system in it
// has a final status.
struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n_sys = 0, n_tiles = 0;
  SysDesc* d_descs = nullptr;
  uint2* d_tiles = nullptr;
  // k_pcg_finish's table: one entry per group of kFinishTiles consecutive
  // tiles of a system (the entries of the tile table whose tile index is a
  // multiple of kFinishTiles); same capacity as d_tiles
  uint2* d_ftiles = nullptr;
  uint32_t n_ftiles = 0;
  SysState* d_states = nullptr;
  SysState* h_init = nullptr;
  SysState* h_poll[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  uint32_t gchunk = 0;
  uint32_t chunk = 16;
  bool use_graph = true;
  uint32_t remap = 1;
  uint32_t update_groups = 1;  // several tiles per k_pcg_update workgroup
  // kernel timing (HIP events around sampled iterations)
  uint32_t timing_every = 0;
  uint64_t trip = 0;  // iterations enqueued so far (for sampling)
  std::vector<hipEvent_t> tev[2];
  std::vector<int> tev_used;  // per pool: number of sampled launch sets recorded
  std::vector<uint64_t> tev_trip[2];  // per pool: each sample's launch set within the run
  // harvested samples of the current run: launch set, then one time per kernel
  std::vector<std::array<double, HBJ_ITER_KERNELS + 1>> t_pending;
  uint64_t run_trip0 = 0;     // `trip` when the current run started
  static constexpr int kKernels = HBJ_ITER_KERNELS;  // spmv, update, finish
  double t_sum[kKernels] = {0, 0, 0};
  uint64_t t_cnt[kKernels] = {0, 0, 0};
  // the last timed run: its samples of launch sets in which some system ran,
  // and each system's launch sets (system i ran in sets [0, sets_last[i]))
  std::vector<std::array<double, HBJ_ITER_KERNELS + 1>> t_last;
  std::vector<uint32_t> sets_last;
  // SysState of system i: d_states[i] (parity 0) and d_states[cap + i]
  // (parity 1); chunks hold an even number of iterations, so copy 0 is
  // current whenever the host looks.
  uint32_t cap = 0;

  hbj_status init(int dev, uint32_t nsys, uint32_t ntiles) {
    device = dev;
    n_sys = nsys;
    n_tiles = ntiles;
    chunk = env_u32("HBJ_CHUNK", 16);
    chunk = chunk < 2 ? 2u : (chunk & ~1u);  // even: the state parity returns to 0
    use_graph = env_u32("HBJ_GRAPH", 1) != 0 && !profiler_tool_attached();
    remap = env_u32("HBJ_XCD_REMAP", 1);
    update_groups = env_u32("HBJ_UPDATE_GROUPS", 1);
    spmv_variant = env_u32("HBJ_SPMV_VARIANT", kSpmvDefault) % kSpmvVariantCount;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    CHK(dalloc(&d_descs, nsys));
    CHK(dalloc(&d_tiles, ntiles));
    CHK(dalloc(&d_ftiles, ntiles));
    tiles_cap = ntiles ? ntiles : 1;
    cap = nsys ? nsys : 1;
    CHK(dalloc(&d_states, 2ull * cap));
    CHK(halloc(&h_init, nsys));
    CHK(halloc(&h_poll[0], nsys));
    CHK(halloc(&h_poll[1], nsys));
    tev_used.assign(2, 0);
    return HBJ_OK;
  }

  // A host-to-device copy that completes before the call returns, on the
  // engine's own (non-blocking) stream: a copy on the legacy stream would be
  // refused -- and would break the capture -- while another engine of the
  // process captures a graph on another thread.
  hbj_status h2d(void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return HBJ_OK;
  }

  // descs[i].st is pointed at this engine's state slot i.
  hbj_status set_table(std::vector<SysDesc> descs, const std::vector<uint2>& tiles) {
    for (uint32_t i = 0; i < descs.size(); ++i) descs[i].st = d_states + i;
    CHK(h2d(d_descs, descs.data(), sizeof(SysDesc) * descs.size()));
    CHK(h2d(d_tiles, tiles.data(), sizeof(uint2) * tiles.size()));
    CHK(upload_groups(tiles));
    n_sys = (uint32_t)descs.size();
    n_tiles = (uint32_t)tiles.size();
    one_mode = n_sys == 1 && env_u32("HBJ_ONE", 1) != 0;
    if (one_mode) h_one = descs[0];
    dot64 = descs.empty() ? 0u : ((descs[0].flags & kFlagDot64) ? 1u : 0u);
    for (const SysDesc& d : descs)
      if (((d.flags & kFlagDot64) ? 1u : 0u) != dot64) {
        set_error("engine: a table cannot mix FP32- and FP64-accumulated systems");
        return HBJ_E_INVALID;
      }
    full_tiles = tiles;
    table_full = true;
    drop_graph();
    return HBJ_OK;
  }

  static std::vector<uint2> groups_of(const std::vector<uint2>& tiles) {
    std::vector<uint2> g;
    g.reserve(tiles.size() / kFinishTiles + 1);
    for (const uint2& t : tiles)
      if (t.y % kFinishTiles == 0) g.push_back(t);
    return g;
  }
  hbj_status upload_groups(const std::vector<uint2>& tiles) {
    const std::vector<uint2> g = groups_of(tiles);
    if (!g.empty()) CHK(h2d(d_ftiles, g.data(), sizeof(uint2) * g.size()));
    n_ftiles = (uint32_t)g.size();
    return HBJ_OK;
  }

  // Tile table compaction for many-system runs: once fewer than 15/16 of the
  // table's tiles belong to running systems, the table is rewritten with only
  // those (finished systems' tiles would otherwise still be dispatched and
  // exit every launch; a compaction drains the stream and recaptures the
  // chunk's graph, and at 15/16 the 256-system batch ran 1.4% faster than at
  // 3/4, 31/32 or on every finished system alike).  `full_tiles` is the table as set; each run starts
  // from it again.
  std::vector<uint2> full_tiles;
  bool table_full = true;

  hbj_status restore_full_table() {
    if (table_full) return HBJ_OK;
    HIPCHK(hipStreamSynchronize(stream));
    CHK(h2d(d_tiles, full_tiles.data(), sizeof(uint2) * full_tiles.size()));
    CHK(upload_groups(full_tiles));
    n_tiles = (uint32_t)full_tiles.size();
    table_full = true;
    drop_graph();
    return HBJ_OK;
  }

  // Returns true when the table was compacted (the stream is idle afterwards
  // and `latest` holds the states of the last enqueued chunk).
  hbj_status maybe_compact(const SysState* polled, const SysState* latest, bool* compacted) {
    *compacted = false;
    if (n_sys < 2 || full_tiles.empty()) return HBJ_OK;
    uint64_t alive = 0;
    for (const uint2& t : full_tiles) alive += polled[t.x].status == kRunning ? 1 : 0;
    if (alive * 16 >= (uint64_t)n_tiles * 15) return HBJ_OK;
    HIPCHK(hipStreamSynchronize(stream));
    std::vector<uint2> live;
    live.reserve(alive);
    for (const uint2& t : full_tiles)
      if (latest[t.x].status == kRunning) live.push_back(t);
    if (live.empty()) return HBJ_OK;  // everything finished in the last chunk
    CHK(h2d(d_tiles, live.data(), sizeof(uint2) * live.size()));
    CHK(upload_groups(live));
    n_tiles = (uint32_t)live.size();
    table_full = false;
    drop_graph();
    if (use_graph && !timing_every) CHK(capture(chunk));
    *compacted = true;
    return HBJ_OK;
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    gchunk = 0;
  }

  uint32_t spmv_variant = kSpmvDefault;
  void set_spmv_variant(uint32_t v) {
    if (v != spmv_variant) {
      spmv_variant = v;
      drop_graph();
    }
  }
  // A table of one system launches the single-system kernels, which take the
  // descriptor as a kernel argument instead of loading it.
  bool one_mode = false;
  SysDesc h_one{};
  // FP64-accumulated dots (kFlagDot64): one setting for the whole table
  uint32_t dot64 = 0;
  LaunchSet ls() const {
    return LaunchSet{d_descs, d_tiles, n_tiles, d_ftiles, n_ftiles, remap, n_sys, spmv_variant,
                     one_mode ? &h_one : nullptr, dot64, 0u, update_groups, cap, nullptr, nullptr};
  }

  // Algorithm of the table's systems: 0 the reference's PCG (three launches
  // per iteration), 1 the one-reduction PCG (two).
  uint32_t alg = 0;
  void set_alg(uint32_t a) {
    if (a != alg) {
      alg = a;
      drop_graph();
    }
  }

  // one PCG iteration (parity par) of every system in the table: the
  // previous finish with this SpMV, update (three-launch form: SpMV,
  // update, finish; one-reduction PCG: SpMV, update)
  bool fused = fused_mode();
  void launch_iter(uint32_t par) {
    LaunchSet l = ls();
    l.par = par;
    if (alg == 1) {
      launch_ccg_spmv(l, stream);
      launch_ccg_update(l, stream);
      return;
    }
    if (fused) {
      launch_pcg_fspmv(l, stream);
      launch_pcg_update(l, stream);
      return;
    }
    launch_pcg_spmv(l, stream);
    launch_pcg_update(l, stream);
    launch_pcg_finish(l, stream);
  }

  hbj_status capture(uint32_t T) {
    drop_graph();
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    for (uint32_t i = 0; i < T; ++i) launch_iter(i & 1u);
    HIPCHK(hipStreamEndCapture(stream, &g));
    hipError_t e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      gexec = nullptr;
      set_error("hipGraphInstantiate failed: %s", hipGetErrorString(e));
      return HBJ_E_DEVICE;
    }
    gchunk = T;
    return HBJ_OK;
  }

  hbj_status ensure_timing_events(uint32_t T) {
    const size_t need = 2u * kKernels * ((T + timing_every - 1) / timing_every + 1);
    for (int b = 0; b < 2; ++b)
      while (tev[b].size() < need) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        tev[b].push_back(e);
      }
    return HBJ_OK;
  }

  hbj_status enqueue_chunk(uint32_t T, int pool) {
    if (timing_every) {
      CHK(ensure_timing_events(T));
      int used = 0;
      tev_trip[pool].clear();
      const LaunchSet l = ls();
      for (uint32_t i = 0; i < T; ++i, ++trip) {
        if ((trip - run_trip0) % timing_every == 0) {
          // a start/stop pair per kernel, set by the kernel's own dispatch
          hipEvent_t* e = &tev[pool][2 * kKernels * used];
          LaunchSet lt = l;
          lt.par = i & 1u;
          auto timed = [&](int k, void (*fn)(const LaunchSet&, hipStream_t)) {
            lt.ev_start = e[2 * k];
            lt.ev_stop = e[2 * k + 1];
            fn(lt, stream);
          };
          if (alg == 1) {  // two kernels: the third sample stays empty
            timed(0, launch_ccg_spmv);
            timed(1, launch_ccg_update);
          } else if (fused) {
            timed(0, launch_pcg_fspmv);
            timed(1, launch_pcg_update);
          } else {
            timed(0, launch_pcg_spmv);
            timed(1, launch_pcg_update);
            timed(2, launch_pcg_finish);
          }
          tev_trip[pool].push_back(trip - run_trip0);
          ++used;
        } else {
          launch_iter(i & 1u);
        }
      }
      tev_used[pool] = used;
      HIPCHK(hipGetLastError());
      return HBJ_OK;
    }
    if (use_graph) {
      if (!gexec || gchunk != T) CHK(capture(T));
      HIPCHK(hipGraphLaunch(gexec, stream));
      return HBJ_OK;
    }
    for (uint32_t i = 0; i < T; ++i) launch_iter(i & 1u);
    HIPCHK(hipGetLastError());
    return HBJ_OK;
  }

  hbj_status harvest_timing(int pool) {
    for (int u = 0; u < tev_used[pool]; ++u) {
      hipEvent_t* e = &tev[pool][2 * kKernels * u];
      std::array<double, kKernels + 1> smp;
      smp[0] = (double)tev_trip[pool][u];
      for (int k = 0; k < kKernels; ++k) {
        smp[k + 1] = -1.0;  // no launch of this kernel (two-launch iterations: no third)
        if ((alg == 1 || fused) && k == 2) continue;
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e[2 * k], e[2 * k + 1]));
        smp[k + 1] = ms * 1000.0;
      }
      t_pending.push_back(smp);
    }
    tev_used[pool] = 0;
    return HBJ_OK;
  }

  // Keeps the samples of launch sets in which some system still ran: a
  // system's `sets` counts the launch sets that found it running, so the
  // run's live sets are [0, max over systems of the sets it added).  Later
  // sampled launches (the rest of the chunk that finished the last system,
  // and the chunk queued behind it) exit at once and would bias the means.
  void commit_timing(const SysState* init, const SysState* out, uint32_t n) {
    uint64_t live = 0;
    sets_last.assign(n, 0u);
    for (uint32_t i = 0; i < n; ++i) {
      const uint64_t s = (uint64_t)(out[i].sets - init[i].sets);
      sets_last[i] = (uint32_t)s;
      live = s > live ? s : live;
    }
    t_last.clear();
    for (const auto& smp : t_pending) {
      if ((uint64_t)smp[0] >= live) continue;
      t_last.push_back(smp);
      for (int k = 0; k < kKernels; ++k) {
        if (smp[k + 1] < 0.0) continue;
        t_sum[k] += smp[k + 1];
        t_cnt[k] += 1;
      }
    }
    t_pending.clear();
  }

  static bool all_done(const SysState* s, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i)
      if (s[i].status == kRunning) return false;
    return true;
  }

  // Host time (ms from the start of the last run's device work) of the
  // status poll that first showed each system's final status.
  std::vector<double> done_ms;
  void note_done(const SysState* polled, double t) {
    for (uint32_t i = 0; i < n_sys; ++i)
      if (done_ms[i] < 0.0 && polled[i].status != kRunning) done_ms[i] = t;
  }

  // Runs until every system has a final status.  `init` holds n_sys states;
  // final states are returned in `out`.
  hbj_status run(const SysState* init, SysState* out, double* ms) {
    done_ms.assign(n_sys, -1.0);
    HIPCHK(hipSetDevice(device));
    if (n_tiles == 0 || all_done(init, n_sys)) {
      std::memcpy(out, init, sizeof(SysState) * n_sys);
      done_ms.assign(n_sys, 0.0);
      if (ms) *ms = 0.0;
      return HBJ_OK;
    }
    // Upper bound on the iterations any system can run (each launch of the
    // three kernels advances every running system by one step): the loop
    // below never polls past it, whatever the device reports.
    uint64_t max_trips = 0;
    for (uint32_t i = 0; i < n_sys; ++i) {
      const bool ir = init[i].phase == kPhIrResid || init[i].phase == kPhIrStart;
      const uint64_t t = ir ? (uint64_t)(init[i].max_outer + 1) * (init[i].max_iter + 4) + 4
                            : (uint64_t)init[i].max_iter + 4;
      max_trips = t > max_trips ? t : max_trips;
    }
    const uint64_t max_chunks = max_trips / chunk + 3;
    CHK(restore_full_table());
    if (use_graph && !timing_every && (!gexec || gchunk != chunk)) {
      HIPCHK(hipStreamSynchronize(stream));
      CHK(capture(chunk));
    }
    std::memcpy(h_init, init, sizeof(SysState) * n_sys);
    HIPCHK(hipMemcpyAsync(d_states, h_init, sizeof(SysState) * n_sys, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_states + cap, h_init, sizeof(SysState) * n_sys, hipMemcpyHostToDevice,
                          stream));
    const double t0 = now_ms();
    uint32_t c = 0;
    int final_pool = -1;
    run_trip0 = trip;
    t_pending.clear();
    for (;;) {
      if (c >= max_chunks) {
        HIPCHK(hipStreamSynchronize(stream));
        set_error("engine: systems still running after %llu iterations (status not advancing)",
                  (unsigned long long)(c * (uint64_t)chunk));
        return HBJ_E_DEVICE;
      }
      const int pool = (int)(c & 1u);
      CHK(enqueue_chunk(chunk, pool));
      HIPCHK(hipMemcpyAsync(h_poll[pool], d_states, sizeof(SysState) * n_sys,
                            hipMemcpyDeviceToHost, stream));
      HIPCHK(hipEventRecord(ev[pool], stream));
      if (c > 0) {
        const int prev = pool ^ 1;
        HIPCHK(hipEventSynchronize(ev[prev]));
        note_done(h_poll[prev], now_ms() - t0);
        if (timing_every) CHK(harvest_timing(prev));
        if (all_done(h_poll[prev], n_sys)) {
          final_pool = pool;
          break;
        }
        bool compacted = false;
        CHK(maybe_compact(h_poll[prev], h_poll[pool], &compacted));
        if (compacted && all_done(h_poll[pool], n_sys)) {
          final_pool = pool;
          break;
        }
      }
      ++c;
    }
    HIPCHK(hipEventSynchronize(ev[final_pool]));
    if (timing_every) CHK(harvest_timing(final_pool));
    const double t_end = now_ms() - t0;
    note_done(h_poll[final_pool], t_end);
    for (double& d : done_ms)
      if (d < 0.0) d = t_end;
    if (ms) *ms = t_end;
    std::memcpy(out, h_poll[final_pool], sizeof(SysState) * n_sys);
    if (timing_every) commit_timing(init, out, n_sys);
    return HBJ_OK;
  }

  // ---- step-wise use (serving): the caller owns the loop and changes the
  // table between chunks without waiting for the device.  Slot and tile
  // writes are copies on the engine stream from pinned staging, so they land
  // after the chunk in flight and before the next one.  The writes made
  // before the chunk of pool p is enqueued use staging half p
  // (stage_begin(p)); the caller starts a half again only after the chunk
  // that followed its previous writes has been polled, or after
  // sync_stream().
  uint32_t tiles_cap = 0;
  struct Staging {
    SysDesc* desc = nullptr;   // cap records
    SysState* st = nullptr;    // cap records
    uint2* tiles = nullptr;    // tiles_cap entries
    uint2* ftiles = nullptr;   // tiles_cap entries (finish groups)
    uint32_t used = 0;         // slot records written since stage_begin
  } stg[2];

  hbj_status stage_begin(int pool) {
    Staging& g = stg[pool];
    if (!g.desc) {
      CHK(halloc(&g.desc, cap));
      CHK(halloc(&g.st, cap));
      CHK(halloc(&g.tiles, tiles_cap));
      CHK(halloc(&g.ftiles, tiles_cap));
    }
    g.used = 0;
    return HBJ_OK;
  }

  // slot <- (d, s); both of the slot's state copies get s.  One write per
  // slot between stage_begin calls.
  hbj_status stage_slot(int pool, uint32_t slot, SysDesc d, const SysState& s) {
    Staging& g = stg[pool];
    if (slot >= cap || g.used >= cap) {
      set_error("engine: slot %u / staging record %u beyond %u slots", slot, g.used, cap);
      return HBJ_E_INVALID;
    }
    if (d.n) {  // a system (not an idle slot): the table's dot accumulation follows it
      const uint32_t want = (d.flags & kFlagDot64) ? 1u : 0u;
      if (want != dot64) {
        dot64 = want;
        drop_graph();
      }
    }
    d.st = d_states + slot;
    const uint32_t k = g.used++;
    g.desc[k] = d;
    g.st[k] = s;
    HIPCHK(hipMemcpyAsync(d_descs + slot, &g.desc[k], sizeof d, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_states + slot, &g.st[k], sizeof s, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_states + cap + slot, &g.st[k], sizeof s, hipMemcpyHostToDevice, stream));
    return HBJ_OK;
  }

  // The tile table of the next chunks.  Growing it waits for the stream (the
  // chunk in flight reads the table) and at least doubles it.
  hbj_status stage_tiles(int pool, const std::vector<uint2>& tiles) {
    const uint32_t need = (uint32_t)tiles.size();
    if (need > tiles_cap) {
      HIPCHK(hipStreamSynchronize(stream));
      const uint32_t nc = need > 2 * tiles_cap ? need : 2 * tiles_cap;
      dfree(d_tiles);
      dfree(d_ftiles);
      CHK(dalloc(&d_tiles, nc));
      CHK(dalloc(&d_ftiles, nc));
      for (Staging& g : stg)
        if (g.tiles) {
          hfree(g.tiles);
          hfree(g.ftiles);
          CHK(halloc(&g.tiles, nc));
          CHK(halloc(&g.ftiles, nc));
        }
      tiles_cap = nc;
    }
    const std::vector<uint2> groups = groups_of(tiles);
    if (need) {
      std::memcpy(stg[pool].tiles, tiles.data(), sizeof(uint2) * need);
      HIPCHK(hipMemcpyAsync(d_tiles, stg[pool].tiles, sizeof(uint2) * need, hipMemcpyHostToDevice,
                            stream));
      std::memcpy(stg[pool].ftiles, groups.data(), sizeof(uint2) * groups.size());
      HIPCHK(hipMemcpyAsync(d_ftiles, stg[pool].ftiles, sizeof(uint2) * groups.size(),
                            hipMemcpyHostToDevice, stream));
    }
    n_tiles = need;
    n_ftiles = (uint32_t)groups.size();
    drop_graph();
    return HBJ_OK;
  }

  hbj_status sync_stream() {
    HIPCHK(hipStreamSynchronize(stream));
    return HBJ_OK;
  }

  // enqueue T iterations of every slot in the table, then an async copy of
  // all slot states to h_poll[pool] and an event
  hbj_status enqueue_poll(uint32_t T, int pool) {
    CHK(enqueue_chunk(T, pool));
    HIPCHK(hipMemcpyAsync(h_poll[pool], d_states, sizeof(SysState) * n_sys, hipMemcpyDeviceToHost,
                          stream));
    HIPCHK(hipEventRecord(ev[pool], stream));
    return HBJ_OK;
  }

  hbj_status wait_poll(int pool) {
    HIPCHK(hipEventSynchronize(ev[pool]));
    return HBJ_OK;
  }

  void destroy() {
    drop_graph();
    for (int b = 0; b < 2; ++b) {
      for (auto e : tev[b]) (void)hipEventDestroy(e);
      tev[b].clear();
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      ev[b] = nullptr;
      hfree(h_poll[b]);
      hfree(stg[b].desc);
      hfree(stg[b].st);
      hfree(stg[b].tiles);
      hfree(stg[b].ftiles);
    }
    hfree(h_init);
    dfree(d_descs);
    dfree(d_tiles);
    dfree(d_ftiles);
    dfree(d_states);
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

// A server's device memory: one allocation of its HBM budget, handed out in
// blocks (FreeRanges).  Taking and returning blocks touch no HIP API, so
// retiring a system never waits for the device, as hipFree does (it returns
// only once all work on the device has finished, other streams' included).
struct DeviceArena {
  char* base = nullptr;
  uint64_t size = 0;
  std::mutex mu;
  FreeRanges ranges;

  hbj_status init(uint64_t bytes) {
    size = FreeRanges::round(bytes);
    CHK(dalloc(&base, size));
    ranges.reset(size);
    return HBJ_OK;
  }
  bool take(uint64_t bytes, uint64_t* off) {
    std::lock_guard<std::mutex> g(mu);
    return ranges.take(bytes, off);
  }
  void give(uint64_t off, uint64_t bytes) {
    std::lock_guard<std::mutex> g(mu);
    ranges.give(off, bytes);
  }
  void destroy() {
    dfree(base);
    ranges.reset(0);
    size = 0;
  }
};

// Offsets of a block's pieces (256-byte aligned, in the order added); assign()
// points each piece's pointer into the block.
struct BlockLayout {
  struct Piece {
    void** ptr;
    uint64_t off;
  };
  std::vector<Piece> pieces;
  uint64_t size = 0;
  template <typename T>
  void add(T** p, uint64_t count) {
    if (count == 0) count = 1;
    size = (size + 255) & ~255ull;
    pieces.push_back(Piece{reinterpret_cast<void**>(p), size});
    size += count * sizeof(T);
  }
  void assign(char* base) const {
    for (const Piece& pc : pieces) *pc.ptr = base + pc.off;
  }
};

}  // namespace hbj

namespace hbj {

// What system creation derives on the host from a validated CSR matrix,
// before it touches the device: the SpMV tables (tile-local lists and entry
// indices, the sliced layout where the system qualifies), the tile headers,
// the per-tile block ranges, and -- for a lean system -- the LU blocks built
// on the host.  A lean system (the server's) keeps on the device only what
// its kernels read: with a sliced layout, no CSR arrays and no per-entry
// indices (the SpMV, its FP64 refinement residual and the block solves
// never read them when the sliced layout runs).
struct SystemImage {
  uint32_t n = 0, nnz = 0, nb = 0, n_tiles = 0, n_starts = 0, flags = 0;
  bool lean = false;  // a sliced system keeps no CSR / lidx on the device
  std::vector<uint16_t> lidx;
  std::vector<uint2> tloc;
  std::vector<uint32_t> tcols, tgrp;
  std::vector<float> sval;
  std::vector<uint32_t> runs;
  std::vector<SliceHdr> shdr;
  std::vector<TileHdr> thdr;
  std::vector<uint2> tile_blk;
  std::vector<float> lu;  // lean systems: host-built LU blocks (36 per block)
  uint32_t local_tiles = 0, sliced_nz_waves = 0;
  uint64_t local_columns = 0, spmv_layout_bytes = 0, spmv_sliced_bytes = 0;
  bool sliced() const { return (flags & kFlagSliced) != 0u; }
  bool no_csr() const { return lean && sliced(); }
};

}  // namespace hbj

// ------------------------------------------------------------------ system

struct hbj_system {
  hbj_ctx* ctx = nullptr;
  uint32_t n = 0, nnz = 0, nb = 0, n_tiles = 0;
  uint32_t flags = 0;
  uint32_t *rp = nullptr, *col = nullptr, *bs = nullptr;
  float *val = nullptr, *lu = nullptr;
  float *x = nullptr, *x0 = nullptr, *b = nullptr, *p = nullptr, *q = nullptr, *z = nullptr;
  float* pn = nullptr;  // p of odd launch sets (the two-launch iteration)
  float *ra = nullptr, *rb = nullptr;
  float* part = nullptr;      // 3 * n_tiles per-tile FP32 partials
  double* part64 = nullptr;  // 5 * n_tiles per-tile FP64 partials
  float* l1 = nullptr;       // 3 * ceil(n_tiles / 256) first-level sums (kFlagL1)
  double* l164 = nullptr;
  double* resid2 = nullptr;  // 2 * n_tiles FP64 partials of the residual check
  double *x64 = nullptr, *x0_64 = nullptr, *b64 = nullptr, *r64 = nullptr, *x64b = nullptr;
  float* b32 = nullptr;
  double* h_part2 = nullptr;  // pinned, 2 * n_tiles (allocated by the first residual check)
  uint2* tile_blk = nullptr;  // per tile: first block touching it, count
  uint16_t* lidx = nullptr;   // tile-local SpMV tables (build_tile_local)
  TileHdr* thdr = nullptr;
  uint32_t* tcols = nullptr;
  uint32_t* tgrp = nullptr;
  float* sval = nullptr;      // sliced SpMV layout (build_sliced; kFlagSliced)
  uint32_t* runs = nullptr;
  SliceHdr* shdr = nullptr;
  uint32_t local_tiles = 0, sliced_nz_waves = 0;
  uint64_t local_columns = 0, spmv_layout_bytes = 0, spmv_sliced_bytes = 0;
  bool lean = false;          // no CSR arrays or entry indices on the device (a served
                              // sliced system: only SpMV variants 4 and 5 may run it)
  int spmv_variant = -1;      // chosen variant (-1: the engine's default)
  uint64_t bytes = 0;
  Engine eng;
  bool eng_ready = false;
  int eng_mode = -1;  // 0 plain FP32, 1 IR
  SysDesc* d_desc1 = nullptr;  // plain descriptor for the FP64 check kernel
  // Every device buffer above is a piece of one block: from hipMalloc, or
  // from a server's arena.  The FP64 refinement buffers are in it when the
  // system was created for refinement, else in a block of their own
  // allocated by the first refinement call.
  char* block = nullptr;
  uint64_t block_bytes = 0;
  DeviceArena* arena = nullptr;
  uint64_t arena_off = 0;
  char* ir_block = nullptr;
  // one-reduction PCG (hbj_system_set_algorithm): its s vectors, a block of
  // their own allocated when the algorithm is chosen
  uint32_t alg = 0;
  float *sa = nullptr, *sb = nullptr;
  char* ccg_block = nullptr;

  // The pieces of the main block for a system of n rows, nnz entries,
  // n_starts block starts, n_tcols tile-local columns and a sliced layout of
  // n_sval values, n_runs runs and n_tgrp column groups (0 when there is none).
  void layout(BlockLayout& L, uint32_t n_, uint32_t nnz_, uint32_t n_starts, uint64_t n_tcols,
              uint64_t n_sval, uint64_t n_runs, uint64_t n_tgrp, bool with_ir) {
    const uint32_t nt = n_ ? (n_ + kTile - 1) / kTile : 1u;
    const uint64_t nb_ = n_starts ? n_starts - 1 : 0;
    L.add(&rp, n_ + 1ull);
    L.add(&col, nnz_ + 4ull);  // 4 entries of padding for 16-byte loads
    L.add(&val, nnz_ + 4ull);
    L.add(&bs, (uint64_t)n_starts);
    L.add(&lu, 36 * (nb_ ? nb_ : 1));
    for (float** v : {&x, &x0, &b, &p, &pn, &q, &z, &ra, &rb}) L.add(v, n_);
    L.add(&part, 3ull * nt);
    L.add(&part64, 5ull * nt);
    L.add(&l1, 3ull * ((nt + 255) / 256));
    L.add(&l164, 3ull * ((nt + 255) / 256));
    L.add(&resid2, 2ull * nt);
    L.add(&d_desc1, 1);
    L.add(&tile_blk, nt);
    L.add(&lidx, nnz_ + 8ull);
    L.add(&thdr, nt);
    L.add(&tcols, n_tcols);
    L.add(&sval, n_sval);
    L.add(&runs, n_runs);
    L.add(&shdr, n_sval ? (uint64_t)nt * kWaves : 0u);
    L.add(&tgrp, n_tgrp);
    if (with_ir) layout_ir(L, n_);
  }
  void layout_ir(BlockLayout& L, uint32_t n_) {
    for (double** v : {&x64, &x0_64, &b64, &r64, &x64b}) L.add(v, n_);
    L.add(&b32, n_);
  }

  // The block of a lean system without CSR (SystemImage::no_csr()), with the
  // refinement buffers: the pieces copied from the host first, up to
  // *data_end, then the ones that start at zero (vectors, partials).  The
  // SpMV never reads row_ptr, col, val or lidx of such a system.
  void layout_lean(BlockLayout& L, const SystemImage& im, uint64_t* data_end) {
    const uint32_t n_ = im.n;
    const uint32_t nt = n_ ? (n_ + kTile - 1) / kTile : 1u;
    const uint64_t nb_ = im.n_starts ? im.n_starts - 1 : 0;
    L.add(&bs, (uint64_t)im.n_starts);
    L.add(&lu, 36 * (nb_ ? nb_ : 1));
    L.add(&tile_blk, nt);
    L.add(&thdr, nt);
    L.add(&tcols, im.tcols.size());
    L.add(&tgrp, im.tgrp.size());
    L.add(&sval, im.sval.size());
    L.add(&runs, im.runs.size());
    L.add(&shdr, im.shdr.size());
    L.add(&b64, n_);
    *data_end = (L.size + 255) & ~255ull;
    for (float** v : {&x, &x0, &b, &p, &pn, &q, &z, &ra, &rb}) L.add(v, n_);
    L.add(&part, 3ull * nt);
    L.add(&part64, 5ull * nt);
    L.add(&l1, 3ull * ((nt + 255) / 256));
    L.add(&l164, 3ull * ((nt + 255) / 256));
    L.add(&resid2, 2ull * nt);
    L.add(&d_desc1, 1);
    for (double** v : {&x64, &x0_64, &r64, &x64b}) L.add(v, n_);
    L.add(&b32, n_);
  }

  // Whether this system's solves use the two-launch iteration: the r.z
  // partials every SpMV workgroup reduces are prefetched with the state up
  // to kFetchMax tiles; past that the reduction waits behind the state, and
  // the three-launch form (its finish kernel reduces in a quarter as many
  // workgroups) measured faster (64^3 lattice: 89.5 vs 92.6 us per iteration).
  bool two_launch() const { return fused_mode() && n_tiles <= kFetchMaxTiles; }

  // l1: the engine launches k_level1 (kFlagL1; a single system's own engine
  // in the three-launch iteration past kFetchMaxTiles tiles)
  SysDesc desc(bool ir, bool fused, bool l1_sums = false) const {
    SysDesc d{};
    d.row_ptr = rp;
    d.col = col;
    d.val = val;
    d.bstart = bs;
    d.lu = lu;
    d.x = x;
    d.q = q;
    d.z = z;
    d.b = ir ? b32 : b;
    d.p = p;
    d.pn = fused ? pn : p;
    d.part = part;
    d.part64 = part64;
    d.x64 = x64;
    d.x64b = x64b;
    d.b64 = b64;
    d.r64 = r64;
    d.r[0] = ra;
    d.r[1] = rb;
    d.s[0] = sa;
    d.s[1] = sb;
    d.tile_blk = tile_blk;
    d.lidx = lidx;
    d.thdr = thdr;
    d.tcols = tcols;
    d.tgrp = tgrp;
    d.sval = sval;
    d.runs = runs;
    d.shdr = shdr;
    d.l1 = l1;
    d.l164 = l164;
    d.st = nullptr;
    d.n = n;
    d.nb = nb;
    d.n_tiles = n_tiles;
    d.flags = flags | (ir ? (kFlagIr | kFlagDot64) : 0u) | (fused ? kFlagTwoLaunch : 0u) |
              (l1_sums ? kFlagL1 : 0u);
    return d;
  }

  // ccg: start the one-reduction PCG (plain solves only)
  SysState init_state(bool ir, uint32_t max_iter, float rel_tol, float abs_tol,
                      const hbj_ir_params* ip, bool ccg, bool fused) const {
    SysState s;
    std::memset(&s, 0, sizeof s);
    s.status = kRunning;
    const bool f = fused;
    s.phase = ir ? (f ? kPhIrStart : kPhIrResid) : (ccg ? kPhCcgInit : (f ? kPhStart : kPhInit));
    s.max_iter = ir ? ip->max_inner : max_iter;
    s.rel_tol2 = ir ? ip->inner_rel_tol * ip->inner_rel_tol : rel_tol * rel_tol;  // compute.rs:256
    s.abs_tol2 = ir ? 0.0f : abs_tol * abs_tol;                                    // compute.rs:257
    if (ir) {
      s.max_outer = ip->max_outer;
      s.target2 = ip->target_rel_res * ip->target_rel_res;
    }
    return s;
  }

  hbj_status ensure_ir_buffers() {
    if (x64) return HBJ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    BlockLayout L;
    layout_ir(L, n);
    CHK(dalloc(&ir_block, L.size));
    bytes += L.size;
    L.assign(ir_block);
    HIPCHK(hipMemset(x0_64, 0, sizeof(double) * n));
    HIPCHK(hipMemset(b64, 0, sizeof(double) * n));
    return HBJ_OK;
  }

  hbj_status ensure_ccg_buffers() {
    if (sa) return HBJ_OK;
    HIPCHK(hipSetDevice(ctx->device));
    BlockLayout L;
    L.add(&sa, n);
    L.add(&sb, n);
    CHK(dalloc(&ccg_block, L.size));
    bytes += L.size;
    L.assign(ccg_block);
    return HBJ_OK;
  }

  // mode 0: plain solves (the system's algorithm), 1: refinement (always
  // the reference's PCG inside)
  hbj_status ensure_engine(int mode) {
    if (!eng_ready) {
      CHK(eng.init(ctx->device, 1, n_tiles));
      eng.fused = two_launch();
      eng_ready = true;
    }
    eng.set_alg(mode == 0 ? alg : 0u);
    if (spmv_variant >= 0) eng.set_spmv_variant((uint32_t)spmv_variant);
    if (eng_mode != mode) {
      std::vector<uint2> tiles(n_tiles);
      for (uint32_t t = 0; t < n_tiles; ++t) tiles[t] = make_uint2(0u, t);
      // first-level sums once per launch for a single system past the
      // prefetch limit (its update and finish workgroups would each reduce
      // every tile partial again)
      const bool l1_sums = !eng.fused && n_tiles > kFetchMaxTiles && env_u32("HBJ_ONE", 1) != 0 &&
                           env_u32("HBJ_L1", 1) != 0;
      CHK(eng.set_table({desc(mode == 1, eng.fused, l1_sums)}, tiles));
      eng_mode = mode;
    }
    return HBJ_OK;
  }

  // ||b - A x|| / ||b|| in FP64 for the FP32 solution in `x` (plain mode).
  hbj_status true_residual(hipStream_t s, double* out) {
    if (lean) {
      set_error("true residual: the system keeps no CSR arrays on the device");
      return HBJ_E_INVALID;
    }
    if (!h_part2) CHK(halloc(&h_part2, 2ull * n_tiles));
    launch_true_resid(d_desc1, x, resid2, n_tiles, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h_part2, resid2, sizeof(double) * 2 * n_tiles, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double rr = 0, bb = 0;
    for (uint32_t t = 0; t < n_tiles; ++t) {
      rr += h_part2[t];
      bb += h_part2[n_tiles + t];
    }
    *out = bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    return HBJ_OK;
  }

  // A system whose block came from an arena and that was never solved on
  // its own (no engine, no residual check, no separate refinement block)
  // returns its block without any HIP call.
  void release() {
    if (eng_ready) eng.destroy();
    eng_ready = false;
    if (arena) {
      arena->give(arena_off, block_bytes);
      arena = nullptr;
      block = nullptr;
    }
    dfree(block);
    dfree(ir_block);
    dfree(ccg_block);
    hfree(h_part2);
  }
};

namespace hbj {
// The host part of system creation for a validated CSR matrix (no HIP
// calls).  lean: build the LU blocks on the host and mark a sliced system
// as keeping no CSR on the device.  Fails only with the LU builder's zero
// pivot (lean) or a system above the per-system size limit.
hbj_status prepare_image(uint32_t n, uint32_t nnz, const uint32_t* row_ptr, const uint32_t* col_idx,
                         const float* values, const uint32_t* block_starts, uint32_t n_starts,
                         bool lean, SystemImage& im);
// Device bytes of a lean system's block with the refinement buffers.
uint64_t lean_block_bytes(const SystemImage& im);
// hbj_system_create with its option: with_ir puts the FP64 refinement
// buffers into the system's block (zeroed).
hbj_status system_create(hbj_ctx* ctx, uint32_t n, uint32_t nnz, const uint32_t* row_ptr,
                         const uint32_t* col_idx, const float* values, const uint32_t* block_starts,
                         uint32_t n_starts, bool with_ir, hbj_system** out);
// A lean system (with the refinement buffers) in a block of `arena` at
// `off` (of lean_block_bytes(im) bytes): the block is written by copies
// enqueued on `stream` from `staging` (pinned, at least the data part of the
// block, free until `done` completes), CSR only when the system is not
// sliced; `done` is recorded after the copies; *copied receives the bytes
// copied from the host.  No host wait.
hbj_status system_place(hbj_ctx* ctx, const SystemImage& im, const uint32_t* row_ptr,
                        const uint32_t* col_idx, const float* values, const uint32_t* block_starts,
                        const double* b64, DeviceArena* arena, uint64_t off, char* staging,
                        hipStream_t stream, hipEvent_t done, hbj_system** out, uint64_t* copied);
}  // namespace hbj
