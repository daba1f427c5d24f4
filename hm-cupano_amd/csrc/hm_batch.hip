// hm_batch.hip -- one pass over frames of mixed pano configurations.
//
// hm-cupano has no counterpart: its apps call process() once per frame, one
// stream operation after another.  Here several pano calls -- rigs of the
// 2-, 3- and N-image classes, each call possibly a batch of frames, with
// different image counts, canvas sizes and level counts -- run as one pass:
//
//   * each call's pipeline is recorded once (hm_batch.hpp): the sequence of
//     kernel bodies, grids and arguments its process() would launch;
//   * launch k of every call forms stage k.  A call's launch depends only on
//     its own earlier launches, so running the stages in order, each to
//     completion, keeps every dependency;
//   * a stage runs as one launch per kernel body it holds (the bodies of one
//     stage run concurrently on forked streams): a persistent kernel whose
//     workgroups walk the work-list of every recorded launch of that body --
//     every tile (x) of every batch entry (z) of every call -- with a stride
//     of the grid size.  The items and their argument blocks travel in the
//     kernel arguments (scalar loads); a list that does not fit is split.
//     One kernel per body keeps each body's own register allocation: a
//     kernel carrying several would run all of them at the occupancy of the
//     largest;
//   * launches a call puts on its side stream (the canvas outside the blend,
//     which nothing in the frame reads) run in a lane beside the stages, from
//     where the call forked to the end of the pass.
//
// The pass is built for fixed buffers (like a captured graph) and can itself
// be captured as a hipGraph and replayed.  The per-pixel arithmetic is the
// bodies' own, so every frame's output is bit-identical to its own process().
#include <string.h>

#include <algorithm>
#include <array>
#include <utility>
#include <memory>
#include <string>
#include <vector>

#include "hm_batch_ops.hpp"
#include "hm_pano.hpp"

namespace hm {
thread_local LaunchRecorder* g_recorder = nullptr;
}  // namespace hm

using namespace hmdev;
using namespace hm;

namespace {

constexpr int kLaneStreams = 3;  // concurrent launches of one stage beside the caller's stream

}  // namespace

struct hmBatch {
  struct Launch {
    StageArgs args;
    int stage, op;
  };
  std::vector<Launch> launches;       // by stage; within a stage, by class
  std::vector<Launch> free_launches;  // side-stream work: stage = the stage after which it may start
  int stages = 0;
  int recorded = 0;  // kernel launches the calls make one by one
  int grid = 0;      // workgroups per launch (0: one per work unit; fewer: a persistent grid)
  int cus = 0;
  // The classes of a stage run concurrently: side streams forked from and
  // joined back into the caller's stream.
  hipStream_t side[kLaneStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kLaneStreams] = {};
  // Work the calls put on their side streams (the canvas outside the blend):
  // nothing of the frame waits for it, so it runs in a lane of its own beside
  // the stages, started where its call forked and joined at the end.
  hipStream_t lane = nullptr;
  hipEvent_t ev_lane_fork = nullptr, ev_lane_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~hmBatch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& st : side)
      if (st) (void)hipStreamDestroy(st);
    if (lane) (void)hipStreamDestroy(lane);
    if (ev_lane_fork) (void)hipEventDestroy(ev_lane_fork);
    if (ev_lane_join) (void)hipEventDestroy(ev_lane_join);
    for (auto& e : ev_join)
      if (e) (void)hipEventDestroy(e);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
  }
  int grid_for(const Launch& l) const {
    return grid > 0 ? (int)std::min<uint32_t>(l.args.total, (uint32_t)grid) : (int)l.args.total;
  }
};

namespace {

hipError_t record_call(const hmBatchCall& c, LaunchRecorder& rec) {
  hmStitchPano* p = c.pano;
  // Everything that allocates or initialises runs first, outside the recording.
  HM_RETURN_IF_ERROR(p->prepare());
  HM_RETURN_IF_ERROR(hipDeviceSynchronize());
  // Launch shapes the batched kernel carries: remap tiles of 4 pixels x 1 or 2
  // rows, no fused remap+down, levels 1 and 0 of the up pass in two launches,
  // one workgroup per canvas tile.  The output is
  // the same for every tuning.
  const hmTuning saved = p->tuning;
  p->tuning.remap_px = 4;
  if (p->tuning.remap_rows != 2) p->tuning.remap_rows = 1;
  p->tuning.fuse_down = 0;
  p->tuning.fuse_up = 0;
  p->tuning.canvas_grid = 0;
  p->apply_tuning();
  g_recorder = &rec;
  const hipError_t e = p->process_any(c.inputs, c.num_inputs, c.canvas, c.fused, nullptr);
  g_recorder = nullptr;
  p->tuning = saved;
  p->apply_tuning();
  if (e != hipSuccess) return e;
  return rec.error;
}

}  // namespace

extern "C" {

int hmBatchCreate(const hmBatchCall* calls, int num_calls, hmBatch** out) {
  if (!out) return hipErrorInvalidValue;
  *out = nullptr;
  if (!calls || num_calls < 1) return hipErrorInvalidValue;
  for (int i = 0; i < num_calls; ++i) {
    if (!calls[i].pano || calls[i].pano->status != hipSuccess) return hipErrorInvalidValue;
    if (calls[i].pano->debug || calls[i].pano->pipeline_f32) return hipErrorNotSupported;
    if (calls[i].pano->served.load()) return hipErrorNotSupported;  // a server's thread launches it
    for (int j = 0; j < i; ++j)  // one object's scratch serves one call of a pass
      if (calls[j].pano == calls[i].pano) return hipErrorInvalidValue;
  }
  std::vector<std::vector<RecordedLaunch>> per_call((size_t)num_calls);  // the caller's stream
  std::vector<RecordedLaunch> free_rec;                                  // side streams
  int recorded = 0;
  for (int i = 0; i < num_calls; ++i) {
    LaunchRecorder rec;
    const hipError_t e = record_call(calls[i], rec);
    if (e != hipSuccess) return e;
    recorded += (int)rec.launches.size();
    for (auto& r : rec.launches) {
      if (r.stream == nullptr)
        per_call[(size_t)i].push_back(std::move(r));
      else
        free_rec.push_back(std::move(r));
    }
  }
  std::unique_ptr<hmBatch> b(new hmBatch);
  b->recorded = recorded;
  int dev = 0;
  HM_RETURN_IF_ERROR(hipGetDevice(&dev));
  HM_RETURN_IF_ERROR(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
  for (auto& st : b->side) HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
  for (auto& e : b->ev_join) HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&b->lane, hipStreamNonBlocking));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&b->ev_lane_fork, hipEventDisableTiming));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&b->ev_lane_join, hipEventDisableTiming));
  size_t max_len = 0;
  for (const auto& v : per_call) max_len = std::max(max_len, v.size());
  for (const auto& r : free_rec) max_len = std::max(max_len, (size_t)r.main_before);
  // Stage k of the staged list: launch k of every call.  Lane group k: the
  // side-stream launches whose call had made k launches when it forked.
  for (int lane = 0; lane < 2; ++lane)
  for (size_t k = 0; k <= max_len; ++k) {
    std::vector<const RecordedLaunch*> group;
    if (lane == 0) {
      for (const auto& v : per_call)
        if (k < v.size()) group.push_back(&v[k]);
    } else {
      for (const auto& r : free_rec)
        if ((size_t)r.main_before == k) group.push_back(&r);
    }
    std::vector<int> ops;  // the bodies of the group, in the carried order
    for_each_batch_op([&](int op) {
      for (const RecordedLaunch* rp : group)
        if (rp->op == op) {
          ops.push_back(op);
          break;
        }
    });
    for (const int op : ops) {
      StageArgs cur{};
      uint32_t bump = 0;  // next free byte of cur.blob
      auto flush = [&]() {
        if (cur.n_items == 0) return;
        (lane == 0 ? b->launches : b->free_launches).push_back(hmBatch::Launch{cur, (int)k, op});
        cur = StageArgs{};
        bump = 0;
      };
      for (const RecordedLaunch* rp : group) {
        if (rp->op != op) continue;
        const RecordedLaunch& r = *rp;
        const uint32_t size = (uint32_t)((r.args.size() + 15) & ~size_t(15));
        if (size > (uint32_t)kStageBlob) return hipErrorNotSupported;
        if (cur.n_items == kStageItems || bump + size > (uint32_t)kStageBlob) flush();
        StageItem& it = cur.items[cur.n_items++];
        it.op = r.op;
        it.off = bump;
        it.first = cur.total;
        it.wgs = r.wgs;
        memcpy(cur.blob + bump, r.args.data(), r.args.size());
        bump += size;
        cur.total += r.wgs * r.batch;
      }
      flush();
    }
  }
  b->stages = 0;
  for (const auto& v : per_call) b->stages = std::max(b->stages, (int)v.size());
  *out = b.release();
  return hipSuccess;
}

int hmBatchSetGrid(hmBatch* batch, int workgroups) {
  if (!batch || workgroups < 0) return hipErrorInvalidValue;
  if (batch->grid != workgroups && batch->exec) {
    (void)hipGraphExecDestroy(batch->exec);
    (void)hipGraphDestroy(batch->graph);
    batch->exec = nullptr;
    batch->graph = nullptr;
  }
  batch->grid = workgroups;
  return hipSuccess;
}

int hmBatchGetInfo(const hmBatch* batch, int* stages, int* launches, int* recorded_launches, int* grid) {
  if (!batch) return hipErrorInvalidValue;
  if (stages) *stages = batch->stages;
  if (launches) *launches = (int)(batch->launches.size() + batch->free_launches.size());
  if (recorded_launches) *recorded_launches = batch->recorded;
  if (grid) {
    int g = batch->grid;
    if (g == 0)
      for (const auto& l : batch->launches) g = std::max(g, batch->grid_for(l));
    *grid = g;
  }
  return hipSuccess;
}

// Stage by stage; a stage's class launches run concurrently (the first on
// the caller's stream, the others on side streams forked from it), and the
// caller's stream waits for all of them before the next stage.  Lane work of
// fork point k is launched on the lane once the stages before k are done;
// the caller's stream waits for the lane at the end.
static hipError_t launch_lane(hmBatch* b, hipStream_t stream, int k, bool* forked) {
  bool any = false;
  for (const auto& l : b->free_launches) any = any || l.stage == k;
  if (!any) return hipSuccess;
  HM_RETURN_IF_ERROR(hipEventRecord(b->ev_lane_fork, stream));
  HM_RETURN_IF_ERROR(hipStreamWaitEvent(b->lane, b->ev_lane_fork, 0));
  for (const auto& l : b->free_launches) {
    if (l.stage != k) continue;
    hipLaunchKernelGGL(batch_kernel(l.op), dim3((unsigned)b->grid_for(l)), dim3(64, 4, 1), 0, b->lane, l.args);
    HM_RETURN_IF_ERROR(hipGetLastError());
  }
  *forked = true;
  return hipSuccess;
}

static hipError_t launch_stages(hmBatch* b, hipStream_t stream) {
  bool lane_used = false;
  int k = 0;
  size_t i = 0;
  while (i < b->launches.size()) {
    const int stage = b->launches[i].stage;
    for (; k <= stage; ++k) HM_RETURN_IF_ERROR(launch_lane(b, stream, k, &lane_used));
    size_t e = i;
    while (e < b->launches.size() && b->launches[e].stage == stage) ++e;
    const int n_side = std::min((int)(e - i) - 1, kLaneStreams);
    if (n_side > 0) {
      HM_RETURN_IF_ERROR(hipEventRecord(b->ev_fork, stream));
      for (int q = 0; q < n_side; ++q) HM_RETURN_IF_ERROR(hipStreamWaitEvent(b->side[q], b->ev_fork, 0));
    }
    for (size_t j = i; j < e; ++j) {
      const int q = (int)(j - i);  // 0: the caller's stream
      hipStream_t st = q == 0 ? stream : b->side[(q - 1) % kLaneStreams];
      const hmBatch::Launch& l = b->launches[j];
      hipLaunchKernelGGL(batch_kernel(l.op), dim3((unsigned)b->grid_for(l)), dim3(64, 4, 1), 0, st, l.args);
      HM_RETURN_IF_ERROR(hipGetLastError());
    }
    for (int q = 0; q < n_side; ++q) {
      HM_RETURN_IF_ERROR(hipEventRecord(b->ev_join[q], b->side[q]));
      HM_RETURN_IF_ERROR(hipStreamWaitEvent(stream, b->ev_join[q], 0));
    }
    i = e;
  }
  for (; k <= b->stages; ++k) HM_RETURN_IF_ERROR(launch_lane(b, stream, k, &lane_used));
  if (lane_used) {
    HM_RETURN_IF_ERROR(hipEventRecord(b->ev_lane_join, b->lane));
    HM_RETURN_IF_ERROR(hipStreamWaitEvent(stream, b->ev_lane_join, 0));
  }
  return hipSuccess;
}

int hmBatchRun(hmBatch* batch, hipStream_t stream, int use_graph) {
  if (!batch) return hipErrorInvalidValue;
  if (!use_graph) return launch_stages(batch, stream);
  if (!batch->exec) {
    hipStream_t cap = nullptr;
    HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t el = launch_stages(batch, cap);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(cap, &g);
      if (el != hipSuccess) e = el;
      if (e == hipSuccess) {
        e = hipGraphInstantiate(&batch->exec, g, nullptr, nullptr, 0);
        if (e == hipSuccess)
          batch->graph = g;
        else
          (void)hipGraphDestroy(g);
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return e;
  }
  return hipGraphLaunch(batch->exec, stream);
}

int hmBatchAutotune(hmBatch* batch, hipStream_t stream, int passes, int* best_grid, float* best_ms) {
  if (!batch || passes < 1) return hipErrorInvalidValue;
  // One workgroup per work unit, or persistent grids of 2 .. 16 workgroups
  // per compute unit; each timed over `passes` passes after one warm-up.
  const int candidates[] = {0, 2, 4, 8, 16};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HM_RETURN_IF_ERROR(hipEventCreate(&e0));
  hipError_t err = hipEventCreate(&e1);
  int best = batch->grid;
  float best_t = 0.0f;
  for (int c : candidates) {
    if (err != hipSuccess) break;
    err = hmBatchSetGrid(batch, c * std::max(batch->cus, 1)) == hipSuccess ? hipSuccess : hipErrorInvalidValue;
    if (err == hipSuccess) err = launch_stages(batch, stream);
    if (err == hipSuccess) err = hipEventRecord(e0, stream);
    for (int k = 0; k < passes && err == hipSuccess; ++k) err = launch_stages(batch, stream);
    if (err == hipSuccess) err = hipEventRecord(e1, stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    if (err == hipSuccess && (best_t == 0.0f || ms < best_t)) {
      best_t = ms;
      best = batch->grid;
    }
  }
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  const hipError_t set = (hipError_t)hmBatchSetGrid(batch, best);
  if (err != hipSuccess) return err;
  if (best_grid) *best_grid = best;
  if (best_ms) *best_ms = best_t / (float)passes;
  return set;
}

int hmBatchDestroy(hmBatch* batch) {
  delete batch;
  return hipSuccess;
}

}  // extern "C"
