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
#include <memory>
#include <string>
#include <vector>

#include "hm_pano.hpp"
#include "hm_pass.hpp"

namespace hm {
thread_local LaunchRecorder* g_recorder = nullptr;

hipError_t record_process(hmStitchPano* p, const hmSurface* inputs, int num_inputs, const hmSurface& canvas,
                          int fused, Recording& out) {
  // Launch shapes the batched kernels carry: remap tiles of 4 pixels x 1 or 2
  // rows, no fused remap+down, one workgroup per canvas tile.  The output is
  // the same for every tuning.
  if (p->debug || p->pipeline_f32) return hipErrorNotSupported;  // launches the batched kernels do not carry
  const hmTuning saved = p->tuning;
  const int status = p->status;  // a refused recording must not become the object's sticky error
  const bool timed = p->timer.enabled();
  p->tuning.remap_px = 4;
  if (p->tuning.remap_rows != 2) p->tuning.remap_rows = 1;
  p->tuning.fuse_down = 0;
  p->tuning.canvas_grid = 0;
  p->apply_tuning();
  p->timer.enable(false);
  LaunchRecorder rec;
  g_recorder = &rec;
  const hipError_t e = p->process_any(inputs, num_inputs, canvas, fused, nullptr);
  g_recorder = nullptr;
  p->timer.enable(timed);
  p->tuning = saved;
  p->apply_tuning();
  p->status = status;
  if (e != hipSuccess) return e;
  if (rec.error != hipSuccess) return rec.error;
  out = std::move(rec.launches);
  return hipSuccess;
}

hipError_t plan_pass(const std::vector<const Recording*>& calls, PassPlan& plan) {
  plan.clear();
  std::vector<std::vector<const RecordedLaunch*>> per_call(calls.size());  // the caller's stream
  std::vector<const RecordedLaunch*> free_rec;                              // side streams
  for (size_t i = 0; i < calls.size(); ++i) {
    plan.recorded += (int)calls[i]->size();
    for (const auto& r : *calls[i]) {
      if (r.stream == nullptr)
        per_call[i].push_back(&r);
      else
        free_rec.push_back(&r);
    }
  }
  size_t max_len = 0;
  for (const auto& v : per_call) max_len = std::max(max_len, v.size());
  for (const auto* r : free_rec) max_len = std::max(max_len, (size_t)r->main_before);
  // Stage k of the staged list: launch k of every call.  Lane group k: the
  // side-stream launches whose call had made k launches when it forked.
  std::vector<const RecordedLaunch*> group;
  for (int lane = 0; lane < 2; ++lane) {
    for (size_t k = 0; k <= max_len; ++k) {
      group.clear();
      if (lane == 0) {
        for (const auto& v : per_call)
          if (k < v.size()) group.push_back(v[k]);
      } else {
        for (const auto* r : free_rec)
          if ((size_t)r->main_before == k) group.push_back(r);
      }
      if (group.empty()) continue;
      std::vector<PassLaunch>& dst = lane == 0 ? plan.launches : plan.free_launches;
      for_each_batch_op([&](int op) {  // the bodies of the group, in the carried order
        bool any = false;
        for (const RecordedLaunch* rp : group) any = any || rp->op == op;
        if (!any) return;
        dst.emplace_back();
        PassLaunch* cur = &dst.back();
        memset(&cur->args, 0, offsetof(StageArgs, blob));
        cur->stage = (int)k;
        cur->op = op;
        uint32_t bump = 0;  // next free byte of the blob
        for (const RecordedLaunch* rp : group) {
          if (rp->op != op) continue;
          const uint32_t size = (uint32_t)((rp->args.size() + 15) & ~size_t(15));
          if (cur->args.n_items == kStageItems || bump + size > (uint32_t)kStageBlob) {
            dst.emplace_back();  // split: a new launch of the same body
            cur = &dst.back();
            memset(&cur->args, 0, offsetof(StageArgs, blob));
            cur->stage = (int)k;
            cur->op = op;
            bump = 0;
          }
          StageItem& it = cur->args.items[cur->args.n_items++];
          it.op = rp->op;
          it.off = bump;
          it.first = cur->args.total;
          it.wgs = rp->wgs;
          memcpy(cur->args.blob + bump, rp->args.data(), rp->args.size());
          bump += size;
          cur->args.total += rp->wgs * rp->batch;
        }
      });
    }
  }
  for (const auto& v : per_call) plan.stages = std::max(plan.stages, (int)v.size());
  return hipSuccess;
}

hipError_t PassStreams::create() {
  if (ready()) return hipSuccess;
  for (auto& st : side) HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  for (auto& e : ev_join) HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&ev_lane_fork, hipEventDisableTiming));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&ev_lane_join, hipEventDisableTiming));
  return hipStreamCreateWithFlags(&lane, hipStreamNonBlocking);
}

void PassStreams::destroy() {
  for (auto& st : side) {
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
  }
  if (lane) (void)hipStreamDestroy(lane);
  lane = nullptr;
  for (hipEvent_t* e : {&ev_fork, &ev_lane_fork, &ev_lane_join, &ev_join[0], &ev_join[1], &ev_join[2]}) {
    if (*e) (void)hipEventDestroy(*e);
    *e = nullptr;
  }
}

// Lane work of fork point k is launched on the lane once the stages before k
// are done.
static hipError_t launch_lane(const PassPlan& plan, PassStreams& ps, int grid, hipStream_t stream, int k,
                              bool* forked) {
  bool any = false;
  for (const auto& l : plan.free_launches) any = any || l.stage == k;
  if (!any) return hipSuccess;
  HM_RETURN_IF_ERROR(hipEventRecord(ps.ev_lane_fork, stream));
  HM_RETURN_IF_ERROR(hipStreamWaitEvent(ps.lane, ps.ev_lane_fork, 0));
  for (const auto& l : plan.free_launches) {
    if (l.stage != k) continue;
    hipLaunchKernelGGL(pass_kernel(l, grid), dim3((unsigned)pass_grid_for(l, grid)), dim3(64, 4, 1), 0, ps.lane, l.args);
    HM_RETURN_IF_ERROR(hipGetLastError());
  }
  *forked = true;
  return hipSuccess;
}

// Stage by stage; a stage's body launches run concurrently (the first on the
// caller's stream, the others on side streams forked from it), and the
// caller's stream waits for all of them before the next stage.  The caller's
// stream waits for the lane at the end.
hipError_t run_pass(const PassPlan& plan, PassStreams& ps, int grid, hipStream_t stream) {
  HM_RETURN_IF_ERROR(ps.create());
  bool lane_used = false;
  int k = 0;
  size_t i = 0;
  while (i < plan.launches.size()) {
    const int stage = plan.launches[i].stage;
    for (; k <= stage; ++k) HM_RETURN_IF_ERROR(launch_lane(plan, ps, grid, stream, k, &lane_used));
    size_t e = i;
    while (e < plan.launches.size() && plan.launches[e].stage == stage) ++e;
    const int n_side = std::min((int)(e - i) - 1, kLaneStreams);
    if (n_side > 0) {
      HM_RETURN_IF_ERROR(hipEventRecord(ps.ev_fork, stream));
      for (int q = 0; q < n_side; ++q) HM_RETURN_IF_ERROR(hipStreamWaitEvent(ps.side[q], ps.ev_fork, 0));
    }
    for (size_t j = i; j < e; ++j) {
      const int q = (int)(j - i);  // 0: the caller's stream
      hipStream_t st = q == 0 ? stream : ps.side[(q - 1) % kLaneStreams];
      const PassLaunch& l = plan.launches[j];
      hipLaunchKernelGGL(pass_kernel(l, grid), dim3((unsigned)pass_grid_for(l, grid)), dim3(64, 4, 1), 0, st, l.args);
      HM_RETURN_IF_ERROR(hipGetLastError());
    }
    for (int q = 0; q < n_side; ++q) {
      HM_RETURN_IF_ERROR(hipEventRecord(ps.ev_join[q], ps.side[q]));
      HM_RETURN_IF_ERROR(hipStreamWaitEvent(stream, ps.ev_join[q], 0));
    }
    i = e;
  }
  for (; k <= plan.stages; ++k) HM_RETURN_IF_ERROR(launch_lane(plan, ps, grid, stream, k, &lane_used));
  if (lane_used) {
    HM_RETURN_IF_ERROR(hipEventRecord(ps.ev_lane_join, ps.lane));
    HM_RETURN_IF_ERROR(hipStreamWaitEvent(stream, ps.ev_lane_join, 0));
  }
  return hipSuccess;
}

}  // namespace hm

using namespace hmdev;
using namespace hm;

struct hmBatch {
  PassPlan plan;
  PassStreams ps;
  int grid = 0;  // workgroups per launch (0: one per work unit; fewer: a persistent grid)
  int cus = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~hmBatch() { drop_graph(); }
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

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
  std::vector<Recording> recs((size_t)num_calls);
  for (int i = 0; i < num_calls; ++i) {
    // Everything that allocates or initialises runs first, outside the recording.
    HM_RETURN_IF_ERROR(calls[i].pano->prepare());
    HM_RETURN_IF_ERROR(hipDeviceSynchronize());
    HM_RETURN_IF_ERROR(record_process(calls[i].pano, calls[i].inputs, calls[i].num_inputs, calls[i].canvas,
                                      calls[i].fused, recs[(size_t)i]));
  }
  std::vector<const Recording*> ptrs;
  for (const auto& r : recs) ptrs.push_back(&r);
  std::unique_ptr<hmBatch> b(new hmBatch);
  HM_RETURN_IF_ERROR(plan_pass(ptrs, b->plan));
  int dev = 0;
  HM_RETURN_IF_ERROR(hipGetDevice(&dev));
  HM_RETURN_IF_ERROR(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev));
  HM_RETURN_IF_ERROR(b->ps.create());
  *out = b.release();
  return hipSuccess;
}

int hmBatchSetGrid(hmBatch* batch, int workgroups) {
  if (!batch || workgroups < 0) return hipErrorInvalidValue;
  if (batch->grid != workgroups) batch->drop_graph();
  batch->grid = workgroups;
  return hipSuccess;
}

int hmBatchGetInfo(const hmBatch* batch, int* stages, int* launches, int* recorded_launches, int* grid) {
  if (!batch) return hipErrorInvalidValue;
  if (stages) *stages = batch->plan.stages;
  if (launches) *launches = (int)(batch->plan.launches.size() + batch->plan.free_launches.size());
  if (recorded_launches) *recorded_launches = batch->plan.recorded;
  if (grid) {
    int g = batch->grid;
    if (g == 0)
      for (const auto& l : batch->plan.launches) g = std::max(g, pass_grid_for(l, batch->grid));
    *grid = g;
  }
  return hipSuccess;
}

int hmBatchRun(hmBatch* batch, hipStream_t stream, int use_graph) {
  if (!batch) return hipErrorInvalidValue;
  if (!use_graph) return run_pass(batch->plan, batch->ps, batch->grid, stream);
  if (!batch->exec) {
    hipStream_t cap = nullptr;
    HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      const hipError_t el = run_pass(batch->plan, batch->ps, batch->grid, cap);
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
    if (err == hipSuccess) err = run_pass(batch->plan, batch->ps, batch->grid, stream);
    if (err == hipSuccess) err = hipEventRecord(e0, stream);
    for (int k = 0; k < passes && err == hipSuccess; ++k) err = run_pass(batch->plan, batch->ps, batch->grid, stream);
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
