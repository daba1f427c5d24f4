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
//   - with batching on (hmServerSetBatching), the requests that are ready
//     when a stream frees up -- frames of different rigs, of any class,
//     canvas size, image and level count -- run as one batched pass: each
//     kernel body of a pipeline stage is one persistent launch whose
//     workgroups walk the tile work-list of every frame in the pass (the
//     pass planner of hm_pass.hpp over each request's recorded launches).
//     A pass whose composition comes back is captured as a hipGraph and
//     replayed.  A request that is ready alone takes its rig's graph;
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
#include "hm_pass.hpp"

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
constexpr size_t kMaxPassGraphs = 64;    // captured batched passes, least recently used evicted
constexpr size_t kMaxPassSightings = 4096;

// One rig's recorded launches for one buffer set (a batched pass's input).
struct Rec {
  uint64_t id = 0;   // unique over the server's life: names the recording in pass keys
  uint64_t sig = 0;  // the pipeline's structure: its bodies in launch order and their streams
  hm::Recording launches;
};

uint64_t signature_of(const hm::Recording& launches) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a over (op, side stream, fork point)
  auto mix = [&](uint64_t v) {
    for (int i = 0; i < 8; ++i) {
      h ^= (v >> (8 * i)) & 0xff;
      h *= 1099511628211ull;
    }
  };
  for (const auto& l : launches) {
    mix((uint64_t)(uint32_t)l.op);
    mix(l.stream == nullptr ? 0u : 1u + (uint64_t)l.main_before);
  }
  return h;
}

struct Rig {
  hmStitchPano* pano = nullptr;
  size_t bytes = 0;
  bool busy = false;   // a request of this rig is on the device
  int queued = 0;      // requests waiting
  bool removing = false;
  bool batchable = true;  // false once a recording failed (a launch the batched kernels do not carry)
  std::map<GraphKey, hipGraphExec_t> graphs;
  std::deque<GraphKey> lru;
  std::map<GraphKey, std::unique_ptr<Rec>> recs;
  std::deque<GraphKey> rec_lru;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Request> reqs;  // one request, or the frames of a batched pass
  int64_t t_start = 0;
  hipGraphExec_t pass_exec = nullptr;  // the captured pass running here, if any (never destroyed meanwhile)
  std::unique_ptr<hm::PassStreams> ps;  // side streams of the batched passes run here
};

// A captured batched pass, keyed by the recordings it runs (in pass order).
using PassKey = std::vector<uint64_t>;

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
  int max_batch = 0;     // frames per batched pass (below 2: no batching)
  int batch_policy = 0;  // 0: any ready requests; 1: only requests whose pipelines have one structure
  bool stop = false;
  int in_flight = 0;
  long graphs_captured = 0, graph_launches = 0, eager_launches = 0;
  long batched_passes = 0, batched_frames = 0, pass_graphs_captured = 0, pass_graph_launches = 0;
  uint64_t next_rec = 1;
  hm::PassPlan plan;  // planning scratch (scheduler thread)
  std::map<PassKey, hipGraphExec_t> pass_graphs;
  std::deque<PassKey> pass_lru;
  std::map<PassKey, int> pass_seen;  // compositions seen once (captured on the second sighting)
  int sticky = hipSuccess;  // first device-side error seen by the scheduler
  std::thread th;

  void run();
  hipError_t launch(Slot& sl, Rig& rig);
  hipError_t launch_pass(Slot& sl, const std::vector<Rec*>& recs);
  Rec* recording_of(Rig& rig, const Request& r);
  void finish(Slot& sl, hipError_t status, int64_t t_done);
  void drop_graphs(Rig& rig) {
    for (auto& kv : rig.graphs) (void)hipGraphExecDestroy(kv.second);
    rig.graphs.clear();
    rig.lru.clear();
    rig.recs.clear();
    rig.rec_lru.clear();
  }
  // Every captured pass (the server is idle).
  void drop_pass_graphs() {
    for (auto& kv : pass_graphs) (void)hipGraphExecDestroy(kv.second);
    pass_graphs.clear();
    pass_lru.clear();
    pass_seen.clear();
  }
  // The captured passes that ran one of `ids` (recordings of an idle rig: none
  // of them is in flight).
  void drop_pass_graphs_of(const std::vector<uint64_t>& ids) {
    for (auto it = pass_graphs.begin(); it != pass_graphs.end();) {
      bool hit = false;
      for (uint64_t id : it->first) hit = hit || std::find(ids.begin(), ids.end(), id) != ids.end();
      if (!hit) {
        ++it;
        continue;
      }
      (void)hipGraphExecDestroy(it->second);
      pass_lru.erase(std::find(pass_lru.begin(), pass_lru.end(), it->first));
      it = pass_graphs.erase(it);
    }
  }
  bool pass_in_flight(hipGraphExec_t ex) const {
    for (const auto& sl : slots)
      if (sl.busy && sl.pass_exec == ex) return true;
    return false;
  }
};

// Starts one request on a free stream: replays the rig's graph for this
// buffer set (capturing it the first time), or launches eagerly.
hipError_t hmServer::launch(Slot& sl, Rig& rig) {
  Request& r = sl.reqs[0];
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

// The rig's recording for the request's buffer set (recorded the first time:
// host work only), or nullptr when the rig's pipeline cannot be batched.
Rec* hmServer::recording_of(Rig& rig, const Request& r) {
  if (!rig.batchable) return nullptr;
  const GraphKey k = key_of(r);
  auto it = rig.recs.find(k);
  if (it != rig.recs.end()) return it->second.get();
  std::unique_ptr<Rec> rec(new Rec);
  if (hm::record_process(rig.pano, r.inputs.data(), (int)r.inputs.size(), r.canvas, r.fused, rec->launches) !=
      hipSuccess) {
    rig.batchable = false;
    return nullptr;
  }
  rec->id = next_rec++;
  rec->sig = signature_of(rec->launches);
  if (rig.recs.size() >= kMaxGraphsPerRig) {
    rig.recs.erase(rig.rec_lru.front());
    rig.rec_lru.pop_front();
  }
  rig.rec_lru.push_back(k);
  return rig.recs.emplace(k, std::move(rec)).first->second.get();
}

// Starts a batched pass over the recordings on the slot's stream: replays
// the pass's graph when this composition was captured, captures it on its
// second sighting, and launches the pass eagerly otherwise.
hipError_t hmServer::launch_pass(Slot& sl, const std::vector<Rec*>& recs) {
  if (!sl.ps) sl.ps.reset(new hm::PassStreams);
  HM_RETURN_IF_ERROR(sl.ps->create());
  PassKey key;
  std::vector<const hm::Recording*> calls;
  for (const Rec* r : recs) {
    key.push_back(r->id);
    calls.push_back(&r->launches);
  }
  ++batched_passes;
  batched_frames += (long)recs.size();
  auto it = pass_graphs.find(key);
  if (it != pass_graphs.end()) {
    ++pass_graph_launches;
    pass_lru.erase(std::find(pass_lru.begin(), pass_lru.end(), key));  // most recently used last
    pass_lru.push_back(key);
    sl.pass_exec = it->second;
    return hipGraphLaunch(it->second, sl.stream);
  }
  HM_RETURN_IF_ERROR(hm::plan_pass(calls, plan));
  bool capture = false;
  if (use_graphs) {
    if (pass_seen.size() >= kMaxPassSightings) pass_seen.clear();
    capture = ++pass_seen[key] >= 2;
  }
  if (!capture) return hm::run_pass(plan, *sl.ps, 0, sl.stream);
  hipGraph_t g = nullptr;
  HM_RETURN_IF_ERROR(hipStreamBeginCapture(sl.stream, hipStreamCaptureModeThreadLocal));
  const hipError_t e = hm::run_pass(plan, *sl.ps, 0, sl.stream);
  const hipError_t ec = hipStreamEndCapture(sl.stream, &g);
  if (e != hipSuccess || ec != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return e != hipSuccess ? e : ec;
  }
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HM_RETURN_IF_ERROR(ei);
  if (pass_graphs.size() >= kMaxPassGraphs) {
    // the least recently used pass that no stream is running
    for (auto k = pass_lru.begin(); k != pass_lru.end(); ++k) {
      auto victim = pass_graphs.find(*k);
      if (pass_in_flight(victim->second)) continue;
      (void)hipGraphExecDestroy(victim->second);
      pass_graphs.erase(victim);
      pass_lru.erase(k);
      break;
    }
  }
  pass_graphs.emplace(key, ex);
  pass_lru.push_back(key);
  pass_seen.erase(key);
  ++pass_graphs_captured;
  ++pass_graph_launches;
  sl.pass_exec = ex;
  return hipGraphLaunch(ex, sl.stream);
}

// Completes every request of the slot with `status`.
void hmServer::finish(Slot& sl, hipError_t status, int64_t t_done) {
  for (const Request& r : sl.reqs) {
    hmServerResult res{};
    res.tag = r.tag;
    res.rig_id = r.rig;
    res.status = status;
    res.t_submit_ns = r.t_submit;
    res.t_start_ns = sl.t_start;
    res.t_done_ns = t_done;
    done.push_back(res);
    auto rit = rigs.find(r.rig);
    if (rit != rigs.end()) rit->second.busy = false;
    --in_flight;
  }
  if (status != hipSuccess && sticky == hipSuccess) sticky = status;
  sl.reqs.clear();
  sl.pass_exec = nullptr;
  sl.busy = false;
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
      finish(sl, q, now_ns());
      progressed = true;
    }
    // 2. admissions, on each free stream: the oldest request whose rig is
    // idle; with batching on, together with the other ready requests of
    // distinct idle rigs (oldest first) that the batched kernels carry.
    for (auto& sl : slots) {
      if (sl.busy) continue;
      auto ready = [&](const Request& r) {
        auto it = rigs.find(r.rig);
        return it != rigs.end() && !it->second.busy;
      };
      auto qi = std::find_if(queue.begin(), queue.end(), ready);
      if (qi == queue.end()) break;
      std::vector<std::deque<Request>::iterator> pick{qi};
      std::vector<Rec*> recs;
      if (max_batch >= 2) {
        Rec* r0 = recording_of(rigs[qi->rig], *qi);
        if (r0) {
          recs.push_back(r0);
          for (auto qj = std::next(qi); qj != queue.end() && (int)pick.size() < max_batch; ++qj) {
            if (!ready(*qj)) continue;
            bool dup = false;
            for (const auto& p : pick) dup = dup || p->rig == qj->rig;
            if (dup) continue;
            Rec* rj = recording_of(rigs[qj->rig], *qj);
            if (!rj || (batch_policy == 1 && rj->sig != r0->sig)) continue;
            pick.push_back(qj);
            recs.push_back(rj);
          }
        }
      }
      sl.reqs.clear();
      std::vector<size_t> pos;  // queue positions of the picked requests, ascending
      for (auto& p : pick) {
        Rig& rig = rigs[p->rig];
        pos.push_back((size_t)(p - queue.begin()));
        sl.reqs.push_back(std::move(*p));
        --rig.queued;
        rig.busy = true;
        ++in_flight;
      }
      // An erase in the middle of a deque invalidates every iterator: erase by
      // position, the last first.
      for (auto p = pos.rbegin(); p != pos.rend(); ++p) queue.erase(queue.begin() + (std::ptrdiff_t)*p);
      sl.busy = true;
      sl.t_start = now_ns();
      hipError_t e = sl.reqs.size() >= 2 ? launch_pass(sl, recs) : launch(sl, rigs[sl.reqs[0].rig]);
      if (e == hipSuccess) e = hipEventRecord(sl.done, sl.stream);
      if (e != hipSuccess) {  // nothing (more) runs for them: complete them with the error now
        (void)hipStreamSynchronize(sl.stream);
        finish(sl, e, sl.t_start);
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
  std::vector<uint64_t> ids;  // the rig's recordings: the passes that ran them go too
  for (const auto& kv : it->second.recs) ids.push_back(kv.second->id);
  sv->drop_graphs(it->second);
  sv->drop_pass_graphs_of(ids);
  sv->used -= it->second.bytes;
  it->second.pano->served.store(false);
  sv->rigs.erase(it);
  return hipSuccess;
}

extern "C" int hmServerSubmitMany(hmServer* sv, const hmServerRequest* reqs, int n) {
  if (!sv || n < 0 || (n > 0 && !reqs)) return hipErrorInvalidValue;
  std::vector<Request> rs((size_t)n);
  for (int i = 0; i < n; ++i) {
    const hmServerRequest& q = reqs[i];
    if (!q.inputs || q.num_inputs < 1 || q.num_inputs > HM_MAX_IMAGES) return hipErrorInvalidValue;
    Request& r = rs[(size_t)i];
    r.rig = q.rig_id;
    r.inputs.assign(q.inputs, q.inputs + q.num_inputs);
    r.canvas = q.canvas;
    r.fused = q.fused;
    r.tag = q.tag;
  }
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (const Request& r : rs) {  // all or nothing
      auto it = sv->rigs.find(r.rig);
      if (it == sv->rigs.end() || it->second.removing) return hipErrorInvalidValue;
      if (it->second.pano->n_images != (int)r.inputs.size()) return hipErrorInvalidValue;
    }
    const int64_t t = now_ns();
    for (Request& r : rs) {
      r.t_submit = t;
      ++sv->rigs[r.rig].queued;
      sv->queue.push_back(std::move(r));
    }
  }
  sv->cv.notify_one();
  return hipSuccess;
}

extern "C" int hmServerSubmit(hmServer* sv, int rig_id, const hmSurface* inputs, int n_inputs, hmSurface canvas,
                              int fused, uint64_t tag) {
  const hmServerRequest r{rig_id, inputs, n_inputs, canvas, fused, tag};
  return hmServerSubmitMany(sv, &r, 1);
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
  st->batched_passes = sv->batched_passes;
  st->batched_frames = sv->batched_frames;
  st->pass_graphs_captured = sv->pass_graphs_captured;
  st->pass_graph_launches = sv->pass_graph_launches;
  return hipSuccess;
}

extern "C" int hmServerSetBatching(hmServer* sv, int max_frames, int policy) {
  if (!sv || max_frames < 0 || policy < 0 || policy > 1) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->max_batch = max_frames;
  sv->batch_policy = policy;
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
  sv->drop_pass_graphs();
  for (auto& sl : sv->slots) {
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
  }
  delete sv;
}
