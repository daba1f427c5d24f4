// hm_frame_loop.hip -- frame ingress/egress around a pano object.
//
// The loop of tests/stitch_two_videos.cpp:552-585 (read BGR frames, upload,
// process, synchronise, download, BGRA -> BGR, write), restructured as a
// pipeline over a ring of pinned host slots and three streams:
//
//   upload stream    host_in[slot] (BGR, 3 B/px) -> dev_raw[slot]
//   caller's stream  [BGR -> BGRA, alpha 255]  process  [BGRA -> BGR] into dev_out[slot]
//   download stream  dev_out[slot] -> host_out[slot]
//
// Each transfer is either a DMA copy (hipMemcpyAsync) or a copy kernel that
// loads from / stores to the mapped pinned buffer directly.
//
// so the upload of frame k+1, the stitch of frame k and the download of
// frame k-1 overlap, and only 3 bytes per pixel cross PCIe each way (the
// reference converts to BGRA on the host and moves 4).  A 3-channel pano
// stitches the BGR frames as they are; a 4-channel one gets them expanded on
// the device and its canvas packed back to BGR there.  One canvas is reused
// for every frame, as the reference's loop does, so the pano's prior-canvas
// semantics are those of a sequential loop; the ordered caller's stream keeps
// frame k+1's stitch behind frame k's packing.
//
// The host thread fills and drains slots through the caller's callbacks; a
// slot is refilled only after its previous frame's download has completed.
#include <string.h>

#include <chrono>
#include <vector>

#include "hm_pano.hpp"

namespace {

using namespace hm;

// BGR (3 B/px) -> BGRA with alpha 255, the cv::cvtColor(BGR2BGRA) of the
// reference's reader; rows of `w` pixels, `rows` rows over the batch.
__global__ void bgr_to_bgra_kernel(const uint8_t* __restrict__ src, size_t src_pitch, uint8_t* __restrict__ dst,
                                   size_t dst_pitch, int w, int rows) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w || y >= rows) return;
  const uint8_t* s = src + (size_t)y * src_pitch + (size_t)x * 3;
  const uint32_t v = (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16) | 0xFF000000u;
  *reinterpret_cast<uint32_t*>(dst + (size_t)y * dst_pitch + (size_t)x * 4) = v;
}

// BGRA -> BGR (cv::COLOR_BGRA2BGR).
__global__ void bgra_to_bgr_kernel(const uint8_t* __restrict__ src, size_t src_pitch, uint8_t* __restrict__ dst,
                                   size_t dst_pitch, int w, int rows) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= w || y >= rows) return;
  const uint32_t v = *reinterpret_cast<const uint32_t*>(src + (size_t)y * src_pitch + (size_t)x * 4);
  uint8_t* d = dst + (size_t)y * dst_pitch + (size_t)x * 3;
  d[0] = (uint8_t)v;
  d[1] = (uint8_t)(v >> 8);
  d[2] = (uint8_t)(v >> 16);
}

// A plain byte copy in 16-byte words (both buffers 16-byte aligned), used to
// move frames across PCIe with loads or stores of mapped pinned memory
// instead of the DMA engines.
__global__ void __launch_bounds__(256) stream_copy_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                          size_t n) {
  const size_t n16 = n / 16;
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    d[i] = s[i];
  if (blockIdx.x == 0)
    for (size_t i = n16 * 16 + threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

constexpr int kCopyBlocks = 512;

hipError_t kernel_copy(const void* src, void* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(stream_copy_kernel, dim3(kCopyBlocks), dim3(256), 0, s, static_cast<const uint8_t*>(src),
                     static_cast<uint8_t*>(dst), n);
  return hipGetLastError();
}

// Pinned, mapped host memory; `dev` is its address for kernels.
struct PinnedBuf {
  void* p = nullptr;
  void* dev = nullptr;
  size_t bytes = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  hipError_t alloc(size_t n) {
    bytes = n;
    HM_RETURN_IF_ERROR(hipHostMalloc(&p, n ? n : 1, hipHostMallocMapped));
    return hipHostGetDevicePointer(&dev, p, 0);
  }
};

struct Events {
  std::vector<hipEvent_t> ev;
  ~Events() {
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  hipError_t make(size_t n, bool timing) {
    ev.assign(n, nullptr);
    for (auto& e : ev) HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming));
    return hipSuccess;
  }
};

struct Stream {
  hipStream_t s = nullptr;
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

struct Slot {
  std::vector<DevBuf> raw;         // per input: uploaded BGR frames
  std::vector<PinnedBuf> host_in;  // per input
  DevBuf out;                      // packed BGR canvas
  PinnedBuf host_out;
  int64_t frame = -1;              // frame in flight in this slot (-1: none)
};

}  // namespace

extern "C" int hmFrameLoopRun(hmStitchPano* pano, int n_inputs, const uint32_t* in_w, const uint32_t* in_h,
                              const hmFrameLoopOptions* options, hmFrameSource source, hmFrameSink sink, void* user,
                              hipStream_t stream, hmFrameLoopStats* stats) {
  hmFrameLoopOptions opt{3, -1, HM_TRANSFER_DEFAULT, HM_TRANSFER_DEFAULT};
  if (options) opt = *options;
  const int slots = opt.slots ? opt.slots : 3;
  const int64_t max_frames = opt.max_frames;
  const bool kernel_up = opt.upload == HM_TRANSFER_KERNEL;
  const bool kernel_down = opt.download == HM_TRANSFER_KERNEL;
  if (!pano || !in_w || !in_h || !source || !sink || n_inputs < 1 || n_inputs > HM_MAX_IMAGES || slots < 1 ||
      slots > 8 || opt.upload < 0 || opt.upload > HM_TRANSFER_KERNEL || opt.download < 0 ||
      opt.download > HM_TRANSFER_KERNEL)
    return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  if (n_inputs != pano->n_images) return hipErrorInvalidValue;
  if (pano->pipeline_f32) return hipErrorNotSupported;  // BGR byte frames feed the uint8 pipeline
  const int C = pano->channels, B = pano->batch;
  const int cw = pano->canvas_w, ch = pano->canvas_h;
  for (int i = 0; i < n_inputs; ++i)
    if (in_w[i] == 0 || in_h[i] == 0) return hipErrorInvalidValue;
  if (stats) memset(stats, 0, sizeof(*stats));

  // ---- resources ----
  Stream up, down;
  HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&up.s, hipStreamNonBlocking));
  HM_RETURN_IF_ERROR(hipStreamCreateWithFlags(&down.s, hipStreamNonBlocking));
  std::vector<Slot> ring(slots);
  std::vector<DevBuf> expanded(C == 4 ? n_inputs : 0);  // BGRA inputs (4-channel pano)
  DevBuf canvas;
  const size_t out_pitch = (size_t)cw * 3, canvas_pitch = (size_t)cw * C;
  const size_t out_bytes = out_pitch * ch * B;
  for (auto& sl : ring) {
    sl.raw = std::vector<DevBuf>(n_inputs);
    sl.host_in = std::vector<PinnedBuf>(n_inputs);
    for (int i = 0; i < n_inputs; ++i) {
      const size_t b = (size_t)in_w[i] * 3 * in_h[i] * B;
      HM_RETURN_IF_ERROR(sl.raw[i].alloc(b));
      HM_RETURN_IF_ERROR(sl.host_in[i].alloc(b));
    }
    HM_RETURN_IF_ERROR(sl.out.alloc(out_bytes));
    HM_RETURN_IF_ERROR(sl.host_out.alloc(out_bytes));
  }
  for (int i = 0; i < (int)expanded.size(); ++i)
    HM_RETURN_IF_ERROR(expanded[i].alloc((size_t)in_w[i] * 4 * in_h[i] * B));
  HM_RETURN_IF_ERROR(canvas.alloc(canvas_pitch * ch * B));
  // A canvas the stitch does not fully cover starts black (the reference
  // allocates its canvas once and leaves it to the first frame).
  HM_RETURN_IF_ERROR(hipMemsetAsync(canvas.p, 0, canvas.bytes, stream));
  HM_RETURN_IF_ERROR(pano->prepare());
  // per slot: upload begin/end, compute begin/end, download begin/end
  Events ev;
  HM_RETURN_IF_ERROR(ev.make((size_t)slots * 6, true));
  auto E = [&](int slot, int k) { return ev.ev[(size_t)slot * 6 + k]; };

  std::vector<void*> host_ptrs(n_inputs);
  std::vector<uint32_t> pitches(n_inputs);
  for (int i = 0; i < n_inputs; ++i) pitches[i] = in_w[i] * 3;
  double up_ms = 0, comp_ms = 0, down_ms = 0;
  int64_t timed = 0;
  int cb_error = 0;
  hipError_t err = hipSuccess;
  const auto t0 = std::chrono::steady_clock::now();
  auto t_first = t0;  // the first frame's delivery

  // Drains slot `s`: waits for its download, times it, hands it to the sink.
  auto drain = [&](int s) -> hipError_t {
    Slot& sl = ring[s];
    if (sl.frame < 0) return hipSuccess;
    HM_RETURN_IF_ERROR(hipEventSynchronize(E(s, 5)));
    float a = 0, b = 0, c = 0;
    if (hipEventElapsedTime(&a, E(s, 0), E(s, 1)) == hipSuccess &&
        hipEventElapsedTime(&b, E(s, 2), E(s, 3)) == hipSuccess &&
        hipEventElapsedTime(&c, E(s, 4), E(s, 5)) == hipSuccess) {
      up_ms += a;
      comp_ms += b;
      down_ms += c;
      ++timed;
    }
    const int rc = sink(user, sl.frame, sl.host_out.p, (uint32_t)cw, (uint32_t)(ch * B), (uint32_t)out_pitch);
    sl.frame = -1;
    if (stats) {
      if (stats->frames == 0) t_first = std::chrono::steady_clock::now();
      stats->frames += 1;
    }
    if (rc != 0) {
      cb_error = rc;
      return hipErrorUnknown;
    }
    return hipSuccess;
  };

  int64_t k = 0;
  for (; max_frames < 0 || k < max_frames; ++k) {
    const int s = (int)(k % slots);
    Slot& sl = ring[s];
    if ((err = drain(s)) != hipSuccess) break;  // the slot's previous frame is out: its buffers are free
    for (int i = 0; i < n_inputs; ++i) host_ptrs[i] = sl.host_in[i].p;
    const int got = source(user, k, host_ptrs.data(), pitches.data());
    if (got == 0) break;  // end of stream
    if (got < 0) {
      cb_error = got;
      err = hipErrorUnknown;
      break;
    }
    // upload
    if ((err = hipEventRecord(E(s, 0), up.s)) != hipSuccess) break;
    for (int i = 0; i < n_inputs && err == hipSuccess; ++i)
      err = kernel_up ? kernel_copy(sl.host_in[i].dev, sl.raw[i].p, sl.raw[i].bytes, up.s)
                      : hipMemcpyAsync(sl.raw[i].p, sl.host_in[i].p, sl.raw[i].bytes, hipMemcpyHostToDevice, up.s);
    if (err != hipSuccess || (err = hipEventRecord(E(s, 1), up.s)) != hipSuccess) break;
    // stitch on the caller's stream
    if ((err = hipStreamWaitEvent(stream, E(s, 1), 0)) != hipSuccess) break;
    if ((err = hipEventRecord(E(s, 2), stream)) != hipSuccess) break;
    hmSurface ins[HM_MAX_IMAGES];
    for (int i = 0; i < n_inputs; ++i) {
      if (C == 4) {
        hipLaunchKernelGGL(bgr_to_bgra_kernel, dim3((in_w[i] + 255) / 256, in_h[i] * B), dim3(256), 0, stream,
                           sl.raw[i].as<uint8_t>(), (size_t)in_w[i] * 3, expanded[i].as<uint8_t>(),
                           (size_t)in_w[i] * 4, (int)in_w[i], (int)(in_h[i] * B));
        ins[i] = hmSurface{expanded[i].p, in_w[i], in_h[i], in_w[i] * 4};
      } else {
        ins[i] = hmSurface{sl.raw[i].p, in_w[i], in_h[i], in_w[i] * 3};
      }
    }
    const hmSurface cs{canvas.p, (uint32_t)cw, (uint32_t)ch, (uint32_t)canvas_pitch};
    if ((err = pano->process_any(ins, n_inputs, cs, 1, stream)) != hipSuccess) break;
    if (C == 4) {
      hipLaunchKernelGGL(bgra_to_bgr_kernel, dim3((cw + 255) / 256, ch * B), dim3(256), 0, stream,
                         canvas.as<uint8_t>(), canvas_pitch, sl.out.as<uint8_t>(), out_pitch, cw, ch * B);
      err = hipGetLastError();
    } else {
      err = hipMemcpyAsync(sl.out.p, canvas.p, out_bytes, hipMemcpyDeviceToDevice, stream);
    }
    if (err != hipSuccess || (err = hipEventRecord(E(s, 3), stream)) != hipSuccess) break;
    // download
    if ((err = hipStreamWaitEvent(down.s, E(s, 3), 0)) != hipSuccess) break;
    if ((err = hipEventRecord(E(s, 4), down.s)) != hipSuccess) break;
    err = kernel_down ? kernel_copy(sl.out.p, sl.host_out.dev, out_bytes, down.s)
                      : hipMemcpyAsync(sl.host_out.p, sl.out.p, out_bytes, hipMemcpyDeviceToHost, down.s);
    if (err != hipSuccess) break;
    if ((err = hipEventRecord(E(s, 5), down.s)) != hipSuccess) break;
    sl.frame = k;
  }
  // drain what is still in flight, oldest first
  for (int64_t j = 0; j < slots; ++j) {
    const int s = (int)((k + j) % slots);
    if (err == hipSuccess) {
      err = drain(s);
    } else if (ring[s].frame >= 0) {
      (void)hipEventSynchronize(E(s, 5));  // never free buffers a copy still uses
      ring[s].frame = -1;
    }
  }
  (void)hipStreamSynchronize(stream);
  (void)hipStreamSynchronize(up.s);
  (void)hipStreamSynchronize(down.s);
  if (stats) {
    const auto t_end = std::chrono::steady_clock::now();
    stats->seconds = std::chrono::duration<double>(t_end - t0).count();
    if (stats->frames > 0) {
      stats->first_frame_seconds = std::chrono::duration<double>(t_first - t0).count();
      stats->steady_seconds = std::chrono::duration<double>(t_end - t_first).count();
      stats->steady_frames = stats->frames - 1;
    }
    if (timed) {
      stats->upload_ms = up_ms / timed;
      stats->stitch_ms = comp_ms / timed;
      stats->download_ms = down_ms / timed;
    }
    stats->callback_error = cb_error;
    stats->bytes_up_per_frame = 0;
    for (int i = 0; i < n_inputs; ++i) stats->bytes_up_per_frame += (uint64_t)in_w[i] * 3 * in_h[i] * B;
    stats->bytes_down_per_frame = out_bytes;
  }
  return err;
}
