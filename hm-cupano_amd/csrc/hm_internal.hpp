// hm_internal.hpp -- host-side pieces shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "hm_pyramid_fused.hpp"

#define HM_RETURN_IF_ERROR(expr)            \
  do {                                      \
    hipError_t _e = (expr);                 \
    if (_e != hipSuccess) return _e;        \
  } while (0)

namespace hm {

// Device allocation owned by a host object, or a view of memory owned
// elsewhere (view(): never freed here).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = false;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  hipError_t alloc(size_t n) {
    reset();
    bytes = n;
    owned = true;
    return hipMalloc(&p, n ? n : 1);
  }
  void view(void* q, size_t n) {
    reset();
    p = q;
    bytes = n;
  }
  void reset() {
    if (p && owned) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    owned = false;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Optional per-launch event timing on the launch stream.  When enabled every
// recorded launch is bracketed by two events; collect() sums the elapsed
// times per kernel name.
class LaunchTimer {
 public:
  ~LaunchTimer();
  void enable(bool on) { enabled_ = on; }
  bool enabled() const { return enabled_; }
  void begin(hipStream_t s);
  void end(hipStream_t s, const char* name);
  // Synchronizes the recorded events and accumulates into totals.
  hipError_t collect();
  void reset_totals() { totals_.clear(); }
  std::string report() const;

 private:
  struct Pending {
    hipEvent_t a, b;
    std::string name;
  };
  bool enabled_ = false;
  hipEvent_t cur_begin_ = nullptr;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  std::map<std::string, std::pair<double, long>> totals_;
  hipEvent_t get_event();
};

// Dispatchers over the kernel templates (hm_pyramid.hip).
hipError_t launch_down(int c, int n, bool in_u8, const void* const* in, float* const* out, int in_w, int in_h,
                       int out_w, int out_h, int batch, hipStream_t s);
hipError_t launch_mask_down(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int m,
                            hipStream_t s);
// Planar normalised weights of the N-image blend from an interleaved mask of
// m channels: plane i = mask_i / (sum of the pixel's values), or mask_i when
// that sum is 0 (the blend's first normalisation, done once per rig).
hipError_t launch_mask_planar(const float* in, int w, int h, float* out, int m, hipStream_t s);

struct UpLaunch {
  const void* g_hi[hmdev::kMaxImages] = {};
  const float* g_lo[hmdev::kMaxImages] = {};
  const float* mask = nullptr;
  const float* r_lo = nullptr;
  float* lap_out[hmdev::kMaxImages] = {};
  float* blend_out = nullptr;
  float* r_out = nullptr;
  uint8_t* u8_out = nullptr;
  size_t u8_pitch = 0;
  int u8_w = 0, u8_h = 0, ox = 0, oy = 0;
  int hi_w = 0, hi_h = 0, lo_w = 0, lo_h = 0;
  int win_x = 0, win_y = 0, win_w = 0, win_h = 0;
};
hipError_t launch_up(int c, int n, int kind, bool hi_u8, const UpLaunch& a, int batch, hipStream_t s);

// Whole-pyramid kernels (hm_pyramid_fused.hpp).  `lv` is a device table:
// lv[j] describes the j-th level produced (down) or level j+1 (up).
// touch: per tile, the images meeting it (three or more images), or nullptr.
hipError_t launch_down_fused(int c, int n, bool in_u8, const void* const* g0, int w0, int h0, int n_out,
                             const hmdev::LevelDesc* lv, const uint32_t* touch, int batch, hipStream_t s);
struct UpQuadLaunch {
  const void* g0[hmdev::kMaxImages] = {};
  const float* g1[hmdev::kMaxImages] = {};
  const void* mask0 = nullptr;
  int mask0_mode = hmdev::kMask0F32;
  const float* r1 = nullptr;  // nullptr: coarsest level, blend only
  int w0 = 0, h0 = 0, w1 = 0, h1 = 0;
  float* f32_out = nullptr;
  uint8_t* u8_out = nullptr;
  size_t u8_pitch = 0;
  int u8_w = 0, u8_h = 0, ox = 0, oy = 0;
  int win_x = 0, win_y = 0, win_w = 0, win_h = 0;
  hmdev::ImgRect z[hmdev::kMaxImages] = {};  // valid region of each image at this level
  const float* wn = nullptr;                 // N-image blend: planar normalised weights (levels >= 1)
  size_t wn_plane = 0;                       // floats per plane
};
hipError_t launch_up_quad(int c, int n, int kind, bool g0_u8, const UpQuadLaunch& u, int batch, hipStream_t s);
// Two images: levels 1 and 0 in one launch (hmdev::up_l10_kernel).  u is the
// level-0 launch; mask1 the level-1 mask; g2/r2 level 2 (nullptr when level 1
// is the coarsest).
hipError_t launch_up_l10(int c, bool g0_u8, const UpQuadLaunch& u, const float* mask1, const float* const* g2,
                         const float* r2, int w2, int h2, int batch, hipStream_t s);

// Three or more images, uint8 level 0 with seam labels: levels 1 and 0 in one
// launch (hmdev::up_l10n_kernel).  u is the level-0 launch; wn1/z1 level 1's
// weights and valid regions; g2/r2 level 2.  Valid only when
// l10n_window_fits holds for u's window.
hipError_t launch_up_l10n(int c, int n, const UpQuadLaunch& u, const float* wn1, size_t wn1_plane,
                          const hmdev::ImgRect* z1, const float* const* g2, const float* r2, int w2, int h2,
                          int batch, hipStream_t s);
// Whether no level-0 tile of the window meets more images (hmdev::l10n_tile_set)
// than the fused launch carries.
bool l10n_window_fits(int n, const hmdev::ImgRect* z0, const hmdev::ImgRect* z1, int win_x, int win_y, int win_w,
                      int win_h);

// The tail of the up pass for two images: levels kTailBase .. levels-1 in
// one launch (hmdev::up_tail_kernel), giving R at kTailBase.
constexpr int kTailBase = 2;
struct UpTailLaunch {
  int n = 0;  // tail levels; [0] = kTailBase
  const float* g[hmdev::kMaxTail][hmdev::kMaxImages] = {};
  hmdev::ImgRect z[hmdev::kMaxTail][hmdev::kMaxImages] = {};  // valid regions (three or more images)
  const float* mask[hmdev::kMaxTail] = {};
  int w[hmdev::kMaxTail] = {}, h[hmdev::kMaxTail] = {};
  float* r_out = nullptr;
  int wx0 = 0, wy0 = 0, wx1 = -1, wy1 = -1;  // base-level pixels produced
};
// Two images (either kind), or three or four (the 3- and N-image blends, in
// the burst form only: u.n - 1 <= hmdev::kTailLds).
hipError_t launch_up_tail(int c, int n, int kind, const UpTailLaunch& u, int batch, hipStream_t s);
constexpr int kMaxTailImages = 4;

// Where the level-0 reconstruction goes.
struct Level0Out {
  float* f32 = nullptr;   // tight float output (blend-context API)
  uint8_t* u8 = nullptr;  // pitched uint8 canvas (pano classes)
  size_t u8_pitch = 0;
  int u8_w = 0, u8_h = 0, ox = 0, oy = 0;
  int win_x = 0, win_y = 0, win_w = -1, win_h = -1;  // level-0 pixels evaluated (-1: all)
};

// A Laplacian-blend pyramid for N images of one geometry.  Levels >= 1 are
// owned; level 0 (images and mask) is supplied by the caller.
class Pyramid {
 public:
  // kind: hmdev::kBlend2 (1-channel mask, level count unclamped),
  //       hmdev::kBlend3 or hmdev::kBlendN (N-channel mask, stop at a side < 2).
  // With an arena, levels >= 1 are views of it (256-byte aligned, in order)
  // and nothing is allocated; hipErrorOutOfMemory when they do not fit
  // (arena_need then holds the bytes they need).  Debug mode needs no arena.
  hipError_t init(int kind, int n, int channels, int width, int height, int levels, int batch, bool debug,
                  void* arena = nullptr, size_t arena_bytes = 0);
  size_t arena_need = 0;
  hipError_t build_mask_pyramid(const float* mask0, hipStream_t s);
  // g0: n level-0 images (uint8 if g0_u8 else float), tight, batch-major.
  hipError_t run(const void* const* g0, bool g0_u8, const Level0Out& out, hipStream_t s, LaunchTimer* timer);
  // Fused path in two halves: Gaussian levels from level `from` on, then the
  // whole reconstruction.
  hipError_t run_down(const void* const* g0, bool g0_u8, int from, hipStream_t s, LaunchTimer* timer);
  hipError_t run_up(const void* const* g0, bool g0_u8, const Level0Out& out, hipStream_t s, LaunchTimer* timer);
  // Sparse images: s0[i] is the level-0 region outside which image i's input
  // is zero (an empty rectangle: zero everywhere).  Derives every level's
  // support and valid region (hmdev::LevelDesc).  Without it (and always in
  // debug mode, and for one or two images) the pyramid is dense.
  hipError_t set_supports(const hmdev::ImgRect* s0);
  int tail_base = kTailBase;  // two images: levels tail_base.. in one launch (0: never)
  bool fuse_up = true;        // levels 1 and 0 in one launch (two images; three or more with labels)
  int l10n_key[4] = {-1, -1, -1, -1};  // the level-0 window l10n_fits was found for
  bool l10n_fits = false;
  hmdev::ImgRect S[hmdev::kMaxLevels][hmdev::kMaxImages] = {};  // supports
  hmdev::ImgRect Z[hmdev::kMaxLevels][hmdev::kMaxImages] = {};  // valid regions

  int kind = 0, n = 0, c = 0, levels = 0, batch = 0;
  bool debug = false;
  int W[hmdev::kMaxLevels] = {}, H[hmdev::kMaxLevels] = {};
  const float* mask0 = nullptr;                     // float level-0 mask (mask pyramid source)
  const void* mask0_fast = nullptr;                 // level-0 mask read by the fused up pass
  int mask0_mode = hmdev::kMask0F32;
  DevBuf g[hmdev::kMaxImages][hmdev::kMaxLevels];  // level >= 1
  DevBuf mask[hmdev::kMaxLevels];                   // level >= 1
  DevBuf wn[hmdev::kMaxLevels];                     // N-image, level >= 1: planar normalised weights
  DevBuf touch[hmdev::kMaxLevels];                  // sparse: per down launch (by first level), tile image sets
  DevBuf r[hmdev::kMaxLevels];                      // level >= 1
  DevBuf lap[hmdev::kMaxImages][hmdev::kMaxLevels]; // debug
  DevBuf blend[hmdev::kMaxLevels];                  // debug
  DevBuf r0_debug;                                  // debug copy of level-0 reconstruction
  hmdev::LevelDesc level_table[hmdev::kMaxLevels] = {};  // levels 1..levels-1 (host; launches copy what they need)
  size_t allocation_size = 0;
  int mask_channels() const { return kind == hmdev::kBlend2 ? 1 : n; }

 private:
  void build_level_table();
};

// Level sizes: the 2-image rule (src/cuda/cudaBlend.cu:860-866) or the
// N/3-image rule that stops when a side < 2 (src/cuda/cudaBlendN.cu:675-691).
int pyramid_sizes(int w, int h, int levels, bool clamp_small_sides, int* widths, int* heights);

// 1-D grid of (tiles_x * tiles_y) workgroups (z = batch) for tile_pixel().
struct TileGrid {
  dim3 grid;
  int tiles_x, tiles_y;
};
inline TileGrid tile_grid(int w, int h, int batch, dim3 block) {
  TileGrid g;
  g.tiles_x = (w + (int)block.x - 1) / (int)block.x;
  g.tiles_y = (h + (int)block.y - 1) / (int)block.y;
  g.grid = dim3((unsigned)(g.tiles_x * g.tiles_y), 1u, (unsigned)batch);
  return g;
}

inline dim3 grid_for(int w, int h, int batch, dim3 block) {
  return dim3((unsigned)((w + (int)block.x - 1) / (int)block.x), (unsigned)((h + (int)block.y - 1) / (int)block.y),
              (unsigned)batch);
}

}  // namespace hm
