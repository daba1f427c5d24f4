// hm_pano_common.hip -- C-ABI shared by every pano object: status, sizes,
// pyramid inspection and dumps, and launch timing.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <filesystem>
#include <string>
#include <vector>

#include "hm_batch.hpp"
#include "hm_files.h"
#include "hm_pano.hpp"

namespace {

__global__ void u8_to_f32_kernel(const uint8_t* in, float* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

}  // namespace

// ===========================================================================
// shared helpers
// ===========================================================================
namespace hm {

hipError_t upload_packed_map(const uint16_t* map_x, const uint16_t* map_y, int w, int h, DevBuf& dst) {
  if (!map_x || !map_y || w < 1 || h < 1) return hipErrorInvalidValue;
  std::vector<uint32_t> packed((size_t)w * h);
  for (size_t i = 0; i < packed.size(); ++i) packed[i] = (uint32_t)map_x[i] | ((uint32_t)map_y[i] << 16);
  HM_RETURN_IF_ERROR(dst.alloc(packed.size() * sizeof(uint32_t)));
  return hipMemcpy(dst.p, packed.data(), packed.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
}

bool pad_replicate(const uint8_t* src, int sw, int sh, int cw, int ch, std::vector<uint8_t>& out) {
  if (!src || sw < 1 || sh < 1 || sw > cw || sh > ch) return false;
  out.resize((size_t)cw * ch);
  for (int y = 0; y < ch; ++y) {
    const uint8_t* row = src + (size_t)std::min(y, sh - 1) * sw;
    for (int x = 0; x < cw; ++x) out[(size_t)y * cw + x] = row[std::min(x, sw - 1)];
  }
  return true;
}

bool surface_ok(const hmSurface& s, int channels, bool need_vec_align) {
  if (!s.d_ptr || s.width == 0 || s.height == 0) return false;
  if (s.pitch < s.width * (uint32_t)channels) return false;
  if (need_vec_align && channels == 4 && ((reinterpret_cast<uintptr_t>(s.d_ptr) & 3u) || (s.pitch & 3u)))
    return false;
  return true;
}

}  // namespace hm

hipError_t hmStitchPano::init_pyramid(int kind, int bw, int bh, const std::vector<float>& mask_host) {
  HM_RETURN_IF_ERROR(pyr.init(kind, n_images, channels, bw, bh, num_levels, batch, debug));
  HM_RETURN_IF_ERROR(mask0.alloc(mask_host.size() * sizeof(float)));
  HM_RETURN_IF_ERROR(hipMemcpy(mask0.p, mask_host.data(), mask_host.size() * sizeof(float), hipMemcpyHostToDevice));
  // The mask pyramid depends only on the seam: built once and kept resident
  // (the reference rebuilds it every frame with the same kernel arithmetic).
  HM_RETURN_IF_ERROR(pyr.build_mask_pyramid(mask0.as<float>(), nullptr));
  HM_RETURN_IF_ERROR(hipStreamSynchronize(nullptr));
  initialized_pyramid = true;
  return hipSuccess;
}


// ===========================================================================
// side stream
// ===========================================================================
hmStitchPano::~hmStitchPano() {
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
}

hipError_t hmStitchPano::ensure_side() {
  if (side) return hipSuccess;
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HM_RETURN_IF_ERROR(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  return hipStreamCreateWithFlags(&side, hipStreamNonBlocking);
}

// While a call is recorded for a batched pass nothing is enqueued: the pass
// orders the side-stream launches itself.
hipError_t hmStitchPano::fork(hipStream_t s) {
  if (hm::g_recorder) return hipSuccess;
  HM_RETURN_IF_ERROR(hipEventRecord(ev_fork, s));
  return hipStreamWaitEvent(side, ev_fork, 0);
}

hipError_t hmStitchPano::join(hipStream_t s) {
  if (hm::g_recorder) return hipSuccess;
  HM_RETURN_IF_ERROR(hipEventRecord(ev_join, side));
  return hipStreamWaitEvent(s, ev_join, 0);
}

extern "C" int hmStitchPanoStatus(const hmStitchPano* pano) { return pano ? pano->status : hipErrorInvalidValue; }

extern "C" int hmStitchPanoPrepare(hmStitchPano* pano) {
  if (!pano) return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  const hipError_t e = pano->prepare();
  if (e != hipSuccess) pano->status = e;
  return e;
}

extern "C" int hmStitchPanoDeviceBytes(const hmStitchPano* pano, size_t* bytes) {
  if (!pano || !bytes) return hipErrorInvalidValue;
  *bytes = pano->device_bytes();
  return hipSuccess;
}

extern "C" int hmStitchPanoGetTuning(const hmStitchPano* pano, hmTuning* tuning) {
  if (!pano || !tuning) return hipErrorInvalidValue;
  *tuning = pano->tuning;
  return hipSuccess;
}

static bool tuning_ok(const hmTuning& t) {
  return (t.remap_px == 2 || t.remap_px == 4 || t.remap_px == 8) && t.tail_base >= 0 && t.tail_base <= 3 &&
         (t.overlap >= 0 && t.overlap <= 3) && (t.remap_xcd == 0 || t.remap_xcd == 1) &&
         (t.remap_rows == 1 || t.remap_rows == 2 || t.remap_rows == 4) &&
         (t.canvas_grid == 0 || t.canvas_grid == 256 || t.canvas_grid == 512 || t.canvas_grid == 1024 ||
          t.canvas_grid == 2048) &&
         (t.fuse_down == 0 || t.fuse_down == 1) && (t.fuse_up == 0 || t.fuse_up == 1);
}

extern "C" int hmStitchPanoSetTuning(hmStitchPano* pano, const hmTuning* tuning) {
  if (!pano || !tuning || !tuning_ok(*tuning)) return hipErrorInvalidValue;
  if (pano->served.load()) return hipErrorNotSupported;
  pano->tuning = *tuning;
  pano->apply_tuning();
  return hipSuccess;
}

extern "C" int hmStitchPanoAutotune(hmStitchPano* pano, const hmSurface* inputs, int num_inputs, hmSurface canvas,
                                    int fused, int frames, hipStream_t stream, hmTuning* best,
                                    float* best_ms_per_frame) {
  if (!pano || !inputs || frames < 1) return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  if (pano->served.load()) return hipErrorNotSupported;
  HM_RETURN_IF_ERROR(pano->prepare());  // allocations and streams first: they decide the knobs
  const hmTuning start = pano->tuning;
  const bool remap = pano->tunes_remap_px();
  std::vector<int> pxs = remap ? std::vector<int>{2, 4, 8} : std::vector<int>{start.remap_px};
  std::vector<int> tbs = pano->tunes_tail() ? std::vector<int>{0, 1, 2, 3} : std::vector<int>{start.tail_base};
  std::vector<int> ovs = pano->tunes_overlap() ? (remap ? std::vector<int>{0, 1, 2, 3} : std::vector<int>{0, 1})
                                                : std::vector<int>{start.overlap};
  std::vector<int> xcs = remap ? std::vector<int>{0, 1} : std::vector<int>{start.remap_xcd};
  std::vector<int> rys = remap ? std::vector<int>{1, 2, 4} : std::vector<int>{start.remap_rows};
  std::vector<int> gds = remap && pano->tunes_overlap() ? std::vector<int>{0, 256, 512, 1024, 2048}
                                                        : std::vector<int>{start.canvas_grid};
  std::vector<int> fds = remap ? std::vector<int>{0, 1} : std::vector<int>{start.fuse_down};
  std::vector<int> fus = pano->tunes_fuse_up() ? std::vector<int>{0, 1} : std::vector<int>{start.fuse_up};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HM_RETURN_IF_ERROR(hipEventCreate(&e0));
  hipError_t err = hipEventCreate(&e1);
  hmTuning chosen = start;
  float chosen_ms = -1.0f;
  std::vector<std::pair<float, hmTuning>> timed;  // every trial, for the confirmation round
  auto time_frames = [&](const hmTuning& t, int n) -> float {
    pano->tuning = t;
    pano->apply_tuning();
    err = pano->process_any(inputs, num_inputs, canvas, fused, stream);  // warm-up
    if (err == hipSuccess) err = hipEventRecord(e0, stream);
    for (int f = 0; f < n && err == hipSuccess; ++f) err = pano->process_any(inputs, num_inputs, canvas, fused, stream);
    if (err == hipSuccess) err = hipEventRecord(e1, stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
    return ms / (float)n;
  };
  auto trial = [&](const hmTuning& t) {
    if (err != hipSuccess) return;
    const float ms = time_frames(t, frames);
    if (err != hipSuccess) return;
    timed.push_back({ms, t});
    if (chosen_ms < 0.0f || ms < chosen_ms) {
      chosen_ms = ms;
      chosen = t;
    }
  };
  // Two stages: the remap's shape (pixels and rows per thread, tile order)
  // first, then the pyramid and stream knobs on top of the best shape.
  for (int px : pxs)
    for (int ry : rys)
      for (int xc : xcs)
        for (int fd : fds)
          trial(hmTuning{px, start.tail_base, start.overlap, xc, ry, start.canvas_grid, fd, start.fuse_up});
  const hmTuning shape = chosen;
  for (int tb : tbs)
    for (int ov : ovs)
      for (int gd : gds)
        for (int fu : fus) {
          if (!ov && gd) continue;  // the grid knob applies to the second stream only
          trial(hmTuning{shape.remap_px, tb, ov, shape.remap_xcd, shape.remap_rows, gd, shape.fuse_down, fu});
        }
  // Confirmation: the four best settings, timed again three times each over
  // four times as many frames, interleaved; the best of the minimum wins (one
  // short trial is noisy next to differences of a few per cent).
  if (err == hipSuccess && timed.size() > 1) {
    std::stable_sort(timed.begin(), timed.end(),
                     [](const std::pair<float, hmTuning>& x, const std::pair<float, hmTuning>& y) { return x.first < y.first; });
    std::vector<hmTuning> cand;
    for (const auto& tt : timed) {
      bool dup = false;
      for (const auto& c : cand) dup = dup || memcmp(&c, &tt.second, sizeof(hmTuning)) == 0;
      if (!dup) cand.push_back(tt.second);
      if (cand.size() == 4) break;
    }
    std::vector<float> best_ms(cand.size(), -1.0f);
    for (int rep = 0; rep < 3 && err == hipSuccess; ++rep)
      for (size_t k = 0; k < cand.size() && err == hipSuccess; ++k) {
        const float ms = time_frames(cand[k], 4 * frames);
        if (err == hipSuccess && (best_ms[k] < 0.0f || ms < best_ms[k])) best_ms[k] = ms;
      }
    if (err == hipSuccess) {
      size_t kb = 0;
      for (size_t k = 1; k < cand.size(); ++k)
        if (best_ms[k] < best_ms[kb]) kb = k;
      chosen = cand[kb];
      chosen_ms = best_ms[kb];
    }
  }
  (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  pano->tuning = err == hipSuccess ? chosen : start;
  pano->apply_tuning();
  if (err != hipSuccess) return err;
  if (best) *best = chosen;
  if (best_ms_per_frame) *best_ms_per_frame = chosen_ms;
  return hipSuccess;
}

extern "C" int hmStitchPanoCanvasSize(const hmStitchPano* pano, int* width, int* height) {
  if (!pano || !width || !height) return hipErrorInvalidValue;
  *width = pano->canvas_w;
  *height = pano->canvas_h;
  return hipSuccess;
}

extern "C" int hmStitchPanoBatchSize(const hmStitchPano* pano) { return pano ? pano->batch : 0; }

extern "C" void hmStitchPanoDestroy(hmStitchPano* pano) { delete pano; }

extern "C" int hmStitchPanoSetDebug(hmStitchPano* pano, int enable) {
  if (!pano) return hipErrorInvalidValue;
  if (pano->initialized_pyramid) return hipErrorNotSupported;  // must precede the first process
  pano->debug = enable != 0;
  return hipSuccess;
}

extern "C" int hmStitchPanoLevels(const hmStitchPano* pano, int* num_levels, int* widths, int* heights) {
  if (!pano || !num_levels) return hipErrorInvalidValue;
  if (!pano->initialized_pyramid) {
    *num_levels = 0;
    return hipSuccess;
  }
  *num_levels = pano->pyr.levels;
  for (int l = 0; l < pano->pyr.levels; ++l) {
    if (widths) widths[l] = pano->pyr.W[l];
    if (heights) heights[l] = pano->pyr.H[l];
  }
  return hipSuccess;
}

extern "C" int hmStitchPanoReadLevel(hmStitchPano* pano, int what, int image, int level, void* host_dst, size_t bytes,
                                     hipStream_t stream) {
  if (!pano || !host_dst || !pano->initialized_pyramid) return hipErrorInvalidValue;
  const hm::Pyramid& p = pano->pyr;
  if (level < 0 || level >= p.levels || image < 0 || image >= p.n) return hipErrorInvalidValue;
  const size_t px = (size_t)p.W[level] * p.H[level];
  const size_t img_floats = px * p.c * p.batch;
  const void* src = nullptr;
  size_t need = img_floats * sizeof(float);
  hm::DevBuf tmp;
  HM_RETURN_IF_ERROR(hipStreamSynchronize(stream));
  switch (what) {
    case HM_LEVEL_GAUSS:
      if (level == 0 && pano->pipeline_f32) {
        src = pano->g0[image].p;  // the float pipeline's full buffer
      } else if (level == 0) {
        HM_RETURN_IF_ERROR(tmp.alloc(need));
        const size_t n = img_floats;
        hipLaunchKernelGGL(u8_to_f32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           pano->g0[image].as<uint8_t>(), tmp.as<float>(), n);
        HM_RETURN_IF_ERROR(hipStreamSynchronize(stream));
        src = tmp.p;
      } else {
        src = p.g[image][level].p;
      }
      break;
    case HM_LEVEL_MASK:
      need = px * p.mask_channels() * sizeof(float);
      src = level ? p.mask[level].p : pano->mask0.p;
      break;
    case HM_LEVEL_LAPLACIAN: src = p.debug ? p.lap[image][level].p : nullptr; break;
    case HM_LEVEL_BLEND: src = p.debug ? p.blend[level].p : nullptr; break;
    case HM_LEVEL_RECONSTRUCT: src = level ? p.r[level].p : (p.debug ? p.r0_debug.p : nullptr); break;
    default: return hipErrorInvalidValue;
  }
  if (!src) return hipErrorNotSupported;
  if (bytes < need) return hipErrorInvalidValue;
  return hipMemcpy(host_dst, src, need, hipMemcpyDeviceToHost);
}

extern "C" int hmStitchPanoSetTiming(hmStitchPano* pano, int enable) {
  if (!pano) return hipErrorInvalidValue;
  pano->timer.enable(enable != 0);
  return hipSuccess;
}

extern "C" int hmStitchPanoTimingReport(hmStitchPano* pano, char* buf, size_t len, int reset) {
  if (!pano || !buf || len == 0) return hipErrorInvalidValue;
  HM_RETURN_IF_ERROR(pano->timer.collect());
  const std::string r = pano->timer.report();
  const size_t n = r.size() < len - 1 ? r.size() : len - 1;
  memcpy(buf, r.data(), n);
  buf[n] = '\0';
  if (reset) pano->timer.reset_totals();
  return hipSuccess;
}

// ===========================================================================
// pyramid dump (CudaStitchPano::dump_soft_blend_pyramid, cudaPano.inl:420-458)
// ===========================================================================
namespace {

hipError_t write_level(const std::string& dir, const std::string& label, const std::string& name, const void* dev,
                       int w, int h, int c) {
  std::error_code ec;
  const std::filesystem::path sub = std::filesystem::path(dir) / label;
  std::filesystem::create_directories(sub, ec);
  if (ec) return hipErrorUnknown;
  std::vector<float> host((size_t)w * h * c);
  HM_RETURN_IF_ERROR(hipMemcpy(host.data(), dev, host.size() * sizeof(float), hipMemcpyDeviceToHost));
  hmRaster r;
  r.data = host.data();
  r.width = (uint32_t)w;
  r.height = (uint32_t)h;
  r.channels = (uint32_t)c;
  r.bits = 32;
  r.sample_format = HM_SAMPLE_FLOAT;
  r.bytes = host.size() * sizeof(float);
  // cv::imwrite's channel convention, so OpenCV-based readers see memory order
  const std::string path = (sub / name).string();
  return hmTiffWrite(path.c_str(), &r, nullptr, 1) == HM_FILE_OK ? hipSuccess : hipErrorUnknown;
}

}  // namespace

extern "C" int hmStitchPanoDumpPyramid(hmStitchPano* pano, const char* directory, hipStream_t stream) {
  if (!pano || !directory) return hipErrorInvalidValue;
  if (pano->status != hipSuccess) return pano->status;
  if (pano->hard_seam || pano->num_levels <= 0 || !pano->initialized_pyramid) return hipErrorInvalidValue;
  HM_RETURN_IF_ERROR(hipStreamSynchronize(stream));
  const hm::Pyramid& p = pano->pyr;
  // Re-evaluate the last frame's pyramid densely from the resident level-0
  // inputs and mask, keeping every level (the fused path keeps none of the
  // Laplacian or blended levels, and the N/3-image Gaussian levels only
  // around each image).  The values are those of the frame's own pyramid.
  hm::Pyramid t;
  HM_RETURN_IF_ERROR(t.init(p.kind, p.n, p.c, p.W[0], p.H[0], pano->num_levels, p.batch, true));
  HM_RETURN_IF_ERROR(t.build_mask_pyramid(pano->mask0.as<float>(), stream));
  const void* g0[hmdev::kMaxImages] = {};
  for (int i = 0; i < p.n; ++i) g0[i] = pano->g0[i].p;
  hm::Level0Out out;
  out.f32 = t.r0_debug.as<float>();
  HM_RETURN_IF_ERROR(t.run(g0, !pano->pipeline_f32, out, stream, nullptr));
  // level-0 Gaussians as float
  hm::DevBuf g0f[hmdev::kMaxImages];
  const size_t n0 = (size_t)t.W[0] * t.H[0] * t.c * t.batch;
  for (int i = 0; i < p.n; ++i) {
    HM_RETURN_IF_ERROR(g0f[i].alloc(n0 * sizeof(float)));
    if (pano->pipeline_f32)
      HM_RETURN_IF_ERROR(hipMemcpyAsync(g0f[i].p, pano->g0[i].p, n0 * sizeof(float), hipMemcpyDeviceToDevice, stream));
    else
      hipLaunchKernelGGL(u8_to_f32_kernel, dim3((unsigned)((n0 + 255) / 256)), dim3(256), 0, stream,
                         pano->g0[i].as<uint8_t>(), g0f[i].as<float>(), n0);
  }
  HM_RETURN_IF_ERROR(hipGetLastError());
  HM_RETURN_IF_ERROR(hipStreamSynchronize(stream));

  const std::string dir = directory;
  std::error_code ec;
  std::filesystem::create_directories(dir, ec);
  if (ec) return hipErrorUnknown;
  {
    FILE* f = fopen((std::filesystem::path(dir) / "metadata.txt").string().c_str(), "w");
    if (!f) return hipErrorUnknown;
    fprintf(f, "num_levels=%d\nbatch_size=%d\nchannels=%d\n", t.levels, t.batch, t.c);
    for (int l = 0; l < t.levels; ++l) fprintf(f, "level_%d=%dx%d\n", l, t.W[l], t.H[l]);
    if (fclose(f) != 0) return hipErrorUnknown;
  }
  const int M = t.mask_channels();
  for (int l = 0; l < t.levels; ++l) {
    const int w = t.W[l], h = t.H[l];
    const size_t img = (size_t)w * h * t.c;  // floats per batch item
    char name[64];
    auto images = [&](const std::string& label, const float* base) -> hipError_t {
      for (int b = 0; b < t.batch; ++b) {
        snprintf(name, sizeof(name), "level_%d_batch_%d.tiff", l, b);
        HM_RETURN_IF_ERROR(write_level(dir, label, name, base + b * img, w, h, t.c));
      }
      return hipSuccess;
    };
    // The reference's order and names (two images: gauss1 gauss2 mask lap1
    // lap2 blend reconstruct); N images number gauss1..gaussN, lap1..lapN.
    for (int i = 0; i < t.n; ++i)
      HM_RETURN_IF_ERROR(images("gauss" + std::to_string(i + 1), l ? t.g[i][l].as<float>() : g0f[i].as<float>()));
    snprintf(name, sizeof(name), "level_%d.tiff", l);
    HM_RETURN_IF_ERROR(write_level(dir, "mask", name, l ? t.mask[l].as<float>() : pano->mask0.as<float>(), w, h, M));
    for (int i = 0; i < t.n; ++i) HM_RETURN_IF_ERROR(images("lap" + std::to_string(i + 1), t.lap[i][l].as<float>()));
    HM_RETURN_IF_ERROR(images("blend", t.blend[l].as<float>()));
    HM_RETURN_IF_ERROR(images("reconstruct", l ? t.r[l].as<float>() : t.r0_debug.as<float>()));
  }
  return hipSuccess;
}
