// hm_server.hip -- the served form of the stitch: many rigs on one GPU.
//
// hm-cupano has no counterpart: its apps call process() once per frame and
// synchronise (tests/stitch_two_videos.cpp:552-585).  Here:
//   - a rig is a pano object (2-, 3- or N-image) whose maps, masks, mask
//     pyramid and pyramid scratch stay resident in HBM; rigs are admitted
//     against a device-memory budget;
//   - frames arrive as requests; a scheduler thread starts each request on
//     one of a fixed set of streams as soon as a stream and the request's rig
//     are free, so requests that arrive while others run are admitted as
//     slots free up and their work overlaps on the device (continuous
//     batching);
//   - a rig's pipeline is captured once per (inputs, canvas) buffer set as a
//     hipGraph and replayed, so a frame costs one graph launch on the host;
//   - completion is observed by polling one event per stream; each result
//     carries its submit, start and completion times.
// Per-pixel work is the pano classes' own, so every frame is bit-identical
// to the same frame stitched alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "hm_pano.hpp"

namespace {

int64_t now_ns() {
  return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Request {
  int rig = -1;
  std::vector<hmSurface> inputs;
  hmSurface canvas{};
  int fused = 1;
  uint64_t tag = 0;
  int64_t t_submit = 0;
};

// Graph cache key: the buffers a captured pipeline is bound to.
struct GraphKey {
  std::vector<uintptr_t> v;
  bool operator<(const GraphKey& o) const { return v < o.v; }
};

GraphKey key_of(const Request& r) {
  GraphKey k;
  for (const auto& s : r.inputs) {
    k.v.push_back(reinterpret_cast<uintptr_t>(s.d_ptr));
    k.v.push_back(((uintptr_t)s.width << 32) | s.height);
    k.v.push_back(s.pitch);
  }
  k.v.push_back(reinterpret_cast<uintptr_t>(r.canvas.d_ptr));
  k.v.push_back(((uintptr_t)r.canvas.width << 32) | r.canvas.height);
  k.v.push_back(r.canvas.pitch);
  k.v.push_back((uintptr_t)r.fused);
  return k;
}

constexpr size_t kMaxGraphsPerRig = 16;

struct Rig {
  hmStitchPano* pano = nullptr;
  size_t bytes = 0;
  bool busy = false;   // a request of this rig is on the device
  int queued = 0;      // requests waiting
  bool removing = false;
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::deque<GraphKey> lru;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  Request req;
  int64_t t_start = 0;
};

}  // namespace

struct hmServer {
  std::mutex mu;
  std::condition_variable cv;       // wakes the scheduler
  std::condition_variable idle_cv;  // wakes waiters
  std::deque<Request> queue;
  std::vector<Slot> slots;
  std::map<int, Rig> rigs;
  int next_rig = 0;
  std::vector<hmServerResult> done;
  size_t budget = 0, used = 0;
  bool use_graphs = true;
  bool stop = false;
  int in_flight = 0;
  long graphs_captured = 0, graph_launches = 0, eager_launches = 0;
  int sticky = hipSuccess;  // first device-side error seen by the scheduler
  std::thread th;

  void run();
  hipError_t launch(Slot& sl, Rig& rig);
  void drop_graphs(Rig& rig) {
    for (auto& kv : rig.graphs) (void)hipGraphExecDestroy(kv.second);
    rig.graphs.clear();
    rig.lru.clear();
  }
};

// Starts one request on a free stream: replays the rig's graph for this
// buffer set (capturing it the first time), or launches eagerly.
hipError_t hmServer::launch(Slot& sl, Rig& rig) {
  Request& r = sl.req;
  if (use_graphs) {
    const GraphKey k = key_of(r);
    auto it = rig.graphs.find(k);
    if (it == rig.graphs.end()) {
      hipGraph_t g = nullptr;
      HM_RETURN_IF_ERROR(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
      const hipError_t e = rig.pano->process_any(r.inputs.data(), (int)r.inputs.size(), r.canvas, r.fused, sl.stream);
      const hipError_t ec = hipStreamEndCapture(sl.stream, &g);
      if (e != hipSuccess || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return e != hipSuccess ? e : ec;
      }
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      HM_RETURN_IF_ERROR(ei);
      if (rig.graphs.size() >= kMaxGraphsPerRig) {
        const GraphKey old = rig.lru.front();
        rig.lru.pop_front();
        (void)hipGraphExecDestroy(rig.graphs[old]);
        rig.graphs.erase(old);
      }
      it = rig.graphs.emplace(k, ex).first;
      rig.lru.push_back(k);
      ++graphs_captured;
    }
    ++graph_launches;
    return hipGraphLaunch(it->second, sl.stream);
  }
  ++eager_launches;
  return rig.pano->process_any(r.inputs.data(), (int)r.inputs.size(), r.canvas, r.fused, sl.stream);
}

void hmServer::run() {
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    bool progressed = false;
    // 1. completions
    for (auto& sl : slots) {
      if (!sl.busy) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      hmServerResult res{};
      res.tag = sl.req.tag;
      res.rig_id = sl.req.rig;
      res.status = q;
      res.t_submit_ns = sl.req.t_submit;
      res.t_start_ns = sl.t_start;
      res.t_done_ns = now_ns();
      if (q != hipSuccess && sticky == hipSuccess) sticky = q;
      done.push_back(res);
      sl.busy = false;
      auto rit = rigs.find(sl.req.rig);
      if (rit != rigs.end()) rit->second.busy = false;
      --in_flight;
      progressed = true;
    }
    // 2. admissions: the oldest request whose rig is idle, on each free stream
    for (auto& sl : slots) {
      if (sl.busy) continue;
      auto qi = std::find_if(queue.begin(), queue.end(), [&](const Request& r) {
        auto it = rigs.find(r.rig);
        return it != rigs.end() && !it->second.busy;
      });
      if (qi == queue.end()) break;
      Rig& rig = rigs[qi->rig];
      sl.req = std::move(*qi);
      queue.erase(qi);
      --rig.queued;
      rig.busy = true;
      sl.busy = true;
      sl.t_start = now_ns();
      ++in_flight;
      hipError_t e = launch(sl, rig);
      if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
      if (e != hipSuccess) {  // nothing (more) runs for it: complete it with the error now
        hmServerResult res{};
        res.tag = sl.req.tag;
        res.rig_id = sl.req.rig;
        res.status = e;
        res.t_submit_ns = sl.req.t_submit;
        res.t_start_ns = res.t_done_ns = sl.t_start;
        done.push_back(res);
        (void)hipStreamSynchronize(sl.stream);
        sl.busy = false;
        rig.busy = false;
        --in_flight;
      }
      progressed = true;
    }
    if (progressed) idle_cv.notify_all();
    if (stop && queue.empty() && in_flight == 0) break;
    if (in_flight > 0) {
      // Poll the device again shortly; new submissions wake us earlier.
      cv.wait_for(lk, std::chrono::microseconds(20));
    } else if (!progressed) {
      cv.wait(lk, [&] { return stop || !queue.empty(); });
    }
  }
}

extern "C" int hmServerCreate(int n_streams, size_t device_budget_bytes, int use_graphs, hmServer** out) {
  if (!out || n_streams < 1 || n_streams > 64) return hipErrorInvalidValue;
  auto* sv = new hmServer();
  sv->budget = device_budget_bytes;
  sv->use_graphs = use_graphs != 0;
  sv->slots.resize(n_streams);
  for (auto& sl : sv->slots) {
    hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
    if (e != hipSuccess) {
      for (auto& s2 : sv->slots) {
        if (s2.done) (void)hipEventDestroy(s2.done);
        if (s2.stream) (void)hipStreamDestroy(s2.stream);
      }
      delete sv;
      return e;
    }
  }
  sv->th = std::thread([sv] { sv->run(); });
  *out = sv;
  return hipSuccess;
}

extern "C" int hmServerAddRig(hmServer* sv, hmStitchPano* pano, int* rig_id) {
  if (!sv || !pano || !rig_id) return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  // Everything the pipeline needs is allocated now, never while serving.
  const hipError_t e = pano->prepare();
  if (e != hipSuccess) return e;
  const size_t bytes = pano->device_bytes();
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->used + bytes > sv->budget) return hipErrorOutOfMemory;
  bool idle = false;
  if (!pano->served.compare_exchange_strong(idle, true)) return hipErrorInvalidValue;  // already served
  Rig r;
  r.pano = pano;
  r.bytes = bytes;
  const int id = sv->next_rig++;
  sv->rigs.emplace(id, std::move(r));
  sv->used += bytes;
  *rig_id = id;
  return hipSuccess;
}

extern "C" int hmServerRemoveRig(hmServer* sv, int rig_id) {
  if (!sv) return hipErrorInvalidValue;
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->rigs.find(rig_id);
  if (it == sv->rigs.end()) return hipErrorInvalidValue;
  it->second.removing = true;
  sv->idle_cv.wait(lk, [&] { return it->second.queued == 0 && !it->second.busy; });
  sv->drop_graphs(it->second);
  sv->used -= it->second.bytes;
  it->second.pano->served.store(false);
  sv->rigs.erase(it);
  return hipSuccess;
}

extern "C" int hmServerSubmit(hmServer* sv, int rig_id, const hmSurface* inputs, int n_inputs, hmSurface canvas,
                              int fused, uint64_t tag) {
  if (!sv || !inputs || n_inputs < 1 || n_inputs > HM_MAX_IMAGES) return hipErrorInvalidValue;
  Request r;
  r.rig = rig_id;
  r.inputs.assign(inputs, inputs + n_inputs);
  r.canvas = canvas;
  r.fused = fused;
  r.tag = tag;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    auto it = sv->rigs.find(rig_id);
    if (it == sv->rigs.end() || it->second.removing) return hipErrorInvalidValue;
    if (it->second.pano->n_images != n_inputs) return hipErrorInvalidValue;
    r.t_submit = now_ns();
    ++it->second.queued;
    sv->queue.push_back(std::move(r));
  }
  sv->cv.notify_one();
  return hipSuccess;
}

extern "C" int hmServerPoll(hmServer* sv, hmServerResult* results, int max_results, int* n_results) {
  if (!sv || !n_results || (max_results > 0 && !results)) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(sv->mu);
  const int n = std::min<int>(max_results, (int)sv->done.size());
  for (int i = 0; i < n; ++i) results[i] = sv->done[i];
  sv->done.erase(sv->done.begin(), sv->done.begin() + n);
  *n_results = n;
  return hipSuccess;
}

extern "C" int hmServerWaitIdle(hmServer* sv) {
  if (!sv) return hipErrorInvalidValue;
  std::unique_lock<std::mutex> lk(sv->mu);
  sv->idle_cv.wait(lk, [&] { return sv->queue.empty() && sv->in_flight == 0; });
  return sv->sticky;
}

extern "C" int hmServerGetStats(hmServer* sv, hmServerStats* st) {
  if (!sv || !st) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(sv->mu);
  st->n_rigs = (int)sv->rigs.size();
  st->device_budget_bytes = sv->budget;
  st->device_used_bytes = sv->used;
  st->queued = (int)sv->queue.size();
  st->in_flight = sv->in_flight;
  st->graphs_captured = sv->graphs_captured;
  st->graph_launches = sv->graph_launches;
  st->eager_launches = sv->eager_launches;
  return hipSuccess;
}

extern "C" void hmServerDestroy(hmServer* sv) {
  if (!sv) return;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->stop = true;
  }
  sv->cv.notify_all();
  if (sv->th.joinable()) sv->th.join();
  for (auto& kv : sv->rigs) {
    sv->drop_graphs(kv.second);
    kv.second.pano->served.store(false);
  }
  for (auto& sl : sv->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete sv;
}
