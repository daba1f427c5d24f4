// hm_pass.hpp -- planning and running a batched pass (host side).
//
// A pass is built from the recorded launches of several pano calls (see
// hm_batch.hpp): launch k of every call forms stage k, each kernel body of a
// stage is one persistent work-list launch, and launches a call put on its
// side stream run in a lane beside the stages.  hmBatch (a fixed pass over
// fixed buffers) and hmServer (one pass per scheduler tick over the requests
// that are ready) both use it.
#pragma once

#include <memory>
#include <vector>

#include "hm_batch_ops.hpp"
#include "hm_stitch.h"

struct hmStitchPano;

namespace hm {

constexpr int kLaneStreams = 3;  // concurrent launches of one stage beside the caller's stream

struct PassLaunch {
  StageArgs args;
  int stage;  // stage index (lane launches: the stage after which they may start)
  int op;
};

struct PassPlan {
  std::vector<PassLaunch> launches;       // by stage; within a stage, in the carried op order
  std::vector<PassLaunch> free_launches;  // side-stream work
  int stages = 0;                         // stages of the longest call
  int recorded = 0;                       // launches the calls make one by one
  void clear() {
    launches.clear();
    free_launches.clear();
    stages = recorded = 0;
  }
};

// One call's recorded launches (arguments bound to the call's buffers).
using Recording = std::vector<RecordedLaunch>;

// Records one process() of a prepared pano object: no allocation and no GPU
// work.  The launch shapes are forced to the ones the batched kernels carry
// for the recording only (the object's tuning is restored); every shape
// computes the same output.
hipError_t record_process(hmStitchPano* p, const hmSurface* inputs, int num_inputs, const hmSurface& canvas,
                          int fused, Recording& out);

// Builds the stages of a pass over the given calls.
hipError_t plan_pass(const std::vector<const Recording*>& calls, PassPlan& plan);

// The streams and events one running pass uses besides the caller's stream.
struct PassStreams {
  hipStream_t side[kLaneStreams] = {};
  hipEvent_t ev_fork = nullptr, ev_join[kLaneStreams] = {};
  hipStream_t lane = nullptr;
  hipEvent_t ev_lane_fork = nullptr, ev_lane_join = nullptr;
  PassStreams() = default;
  PassStreams(const PassStreams&) = delete;
  PassStreams& operator=(const PassStreams&) = delete;
  ~PassStreams() { destroy(); }
  hipError_t create();
  void destroy();
  bool ready() const { return lane != nullptr; }
};

// Workgroups of a launch: one per work unit (grid 0), or at most `grid`.
inline int pass_grid_for(const PassLaunch& l, int grid) {
  return grid > 0 ? (int)(l.args.total < (uint32_t)grid ? l.args.total : (uint32_t)grid) : (int)l.args.total;
}

// The kernel of a launch: the loop-free form when the grid has a workgroup
// per work unit, the list-walking form for a smaller (persistent) grid.
inline StageKernel pass_kernel(const PassLaunch& l, int grid) {
  return batch_kernel(l.op, grid > 0 && (uint32_t)grid < l.args.total);
}

// Enqueues the pass on `stream` (stage by stage, the bodies of a stage on
// forked streams, lane work beside them, everything joined back into
// `stream`).  Capturable.
hipError_t run_pass(const PassPlan& plan, PassStreams& ps, int grid, hipStream_t stream);

}  // namespace hm
