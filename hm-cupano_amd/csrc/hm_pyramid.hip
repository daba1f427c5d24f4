// hm_pyramid.hip -- pyramid engine: kernel dispatch, level management, the
// kernel-tier blend-context C-ABI and launch timing.
#include <stdio.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <sstream>
#include <type_traits>
#include <vector>

#include "hm_batch.hpp"
#include "hm_internal.hpp"
#include "hm_stitch.h"

namespace hm {

using namespace hmdev;

static const dim3 kBlock(64, 4, 1);  // one 64-lane wavefront per row segment

int pyramid_sizes(int w, int h, int levels, bool clamp_small_sides, int* widths, int* heights) {
  if (levels < 1 || levels > kMaxLevels || w < 1 || h < 1) return -1;
  widths[0] = w;
  heights[0] = h;
  for (int i = 1; i < levels; ++i) {
    if (clamp_small_sides && (widths[i - 1] < 2 || heights[i - 1] < 2)) return i;
    widths[i] = (widths[i - 1] + 1) / 2;
    heights[i] = (heights[i - 1] + 1) / 2;
  }
  return levels;
}

// Mask pyramid: plain mean of the in-range taps over M interleaved channels
// (cudaBlend.cu:268-296; cudaBlendN.cu:19-48; cudaBlend3.cu:220-262).
static __global__ void __launch_bounds__(256) mask_down_kernel(const float* in, int in_w, int in_h, float* out, int out_w,
                                                        int out_h, int m) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= out_w || y >= out_h) return;
  float sums[kMaxImages];
  for (int k = 0; k < m; ++k) sums[k] = 0.0f;
  int count = 0;
  for (int dy = 0; dy < 2; ++dy)
    for (int dx = 0; dx < 2; ++dx) {
      const int ix = 2 * x + dx, iy = 2 * y + dy;
      if (ix < in_w && iy < in_h) {
        const float* p = in + ((size_t)iy * in_w + ix) * m;
        for (int k = 0; k < m; ++k) sums[k] += p[k];
        ++count;
      }
    }
  float* o = out + ((size_t)y * out_w + x) * m;
  for (int k = 0; k < m; ++k) o[k] = sums[k] / (float)count;
}

// The N-image blend's first normalisation of every pixel, stored as planes.
template <int M>
static __global__ void __launch_bounds__(256) mask_planar_kernel(const float* in, int w, int h, float* out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= w || y >= h) return;
  const size_t o = (size_t)y * w + x, plane = (size_t)w * h;
  float v[M];
  blendN_normalize<M>(in + o * M, v);
#pragma unroll
  for (int k = 0; k < M; ++k) out[k * plane + o] = v[k];
}

// ---------------------------------------------------------------------------
// dispatch
// ---------------------------------------------------------------------------
template <typename TIn, int C, int N>
static hipError_t down_t(const void* const* in, float* const* out, int in_w, int in_h, int out_w, int out_h,
                         int batch, hipStream_t s) {
  DownArgs<TIn, N> a;
  for (int i = 0; i < N; ++i) {
    a.in[i] = static_cast<const TIn*>(in[i]);
    a.out[i] = out[i];
  }
  a.in_w = in_w;
  a.in_h = in_h;
  a.out_w = out_w;
  a.out_h = out_h;
  const TileGrid tg = tile_grid(out_w, out_h, batch, kBlock);
  a.tiles_x = tg.tiles_x;
  a.tiles_y = tg.tiles_y;
  if (g_recorder) return g_recorder->fail("down");
  hipLaunchKernelGGL((down_kernel<TIn, C, N>), tg.grid, kBlock, 0, s, a);
  return hipGetLastError();
}

template <typename TIn, int C>
static hipError_t down_c(int n, const void* const* in, float* const* out, int in_w, int in_h, int out_w, int out_h,
                         int batch, hipStream_t s) {
  switch (n) {
    case 2: return down_t<TIn, C, 2>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 3: return down_t<TIn, C, 3>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 4: return down_t<TIn, C, 4>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 5: return down_t<TIn, C, 5>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 6: return down_t<TIn, C, 6>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 7: return down_t<TIn, C, 7>(in, out, in_w, in_h, out_w, out_h, batch, s);
    case 8: return down_t<TIn, C, 8>(in, out, in_w, in_h, out_w, out_h, batch, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_down(int c, int n, bool in_u8, const void* const* in, float* const* out, int in_w, int in_h,
                       int out_w, int out_h, int batch, hipStream_t s) {
  if (c == 4) return in_u8 ? down_c<uint8_t, 4>(n, in, out, in_w, in_h, out_w, out_h, batch, s)
                           : down_c<float, 4>(n, in, out, in_w, in_h, out_w, out_h, batch, s);
  if (c == 3) return in_u8 ? down_c<uint8_t, 3>(n, in, out, in_w, in_h, out_w, out_h, batch, s)
                           : down_c<float, 3>(n, in, out, in_w, in_h, out_w, out_h, batch, s);
  return hipErrorInvalidValue;
}

hipError_t launch_mask_down(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int m,
                            hipStream_t s) {
  if (m < 1 || m > kMaxImages) return hipErrorInvalidValue;
  if (g_recorder) return g_recorder->fail("mask_down");
  hipLaunchKernelGGL(mask_down_kernel, grid_for(out_w, out_h, 1, kBlock), kBlock, 0, s, in, in_w, in_h, out, out_w,
                     out_h, m);
  return hipGetLastError();
}

hipError_t launch_mask_planar(const float* in, int w, int h, float* out, int m, hipStream_t s) {
  const dim3 grid = grid_for(w, h, 1, kBlock);
  if (g_recorder) return g_recorder->fail("mask_planar");
  switch (m) {
    case 3: hipLaunchKernelGGL(mask_planar_kernel<3>, grid, kBlock, 0, s, in, w, h, out); break;
    case 4: hipLaunchKernelGGL(mask_planar_kernel<4>, grid, kBlock, 0, s, in, w, h, out); break;
    case 5: hipLaunchKernelGGL(mask_planar_kernel<5>, grid, kBlock, 0, s, in, w, h, out); break;
    case 6: hipLaunchKernelGGL(mask_planar_kernel<6>, grid, kBlock, 0, s, in, w, h, out); break;
    case 7: hipLaunchKernelGGL(mask_planar_kernel<7>, grid, kBlock, 0, s, in, w, h, out); break;
    case 8: hipLaunchKernelGGL(mask_planar_kernel<8>, grid, kBlock, 0, s, in, w, h, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename THi, int C, int N, int KIND>
static hipError_t up_t(const UpLaunch& u, int batch, hipStream_t s) {
  UpArgs<THi, N> a;
  for (int i = 0; i < N; ++i) {
    a.g_hi[i] = static_cast<const THi*>(u.g_hi[i]);
    a.g_lo[i] = u.g_lo[i];
    a.lap_out[i] = u.lap_out[i];
  }
  a.mask = u.mask;
  a.r_lo = u.r_lo;
  a.blend_out = u.blend_out;
  a.r_out = u.r_out;
  a.u8_out = u.u8_out;
  a.u8_pitch = u.u8_pitch;
  a.u8_w = u.u8_w;
  a.u8_h = u.u8_h;
  a.ox = u.ox;
  a.oy = u.oy;
  a.hi_w = u.hi_w;
  a.hi_h = u.hi_h;
  a.lo_w = u.lo_w;
  a.lo_h = u.lo_h;
  a.win_x = u.win_x;
  a.win_y = u.win_y;
  a.win_w = u.win_w;
  a.win_h = u.win_h;
  if (u.win_w <= 0 || u.win_h <= 0) return hipSuccess;
  const TileGrid tg = tile_grid(u.win_w, u.win_h, batch, kBlock);
  a.tiles_x = tg.tiles_x;
  a.tiles_y = tg.tiles_y;
  if (g_recorder) return g_recorder->fail("up");
  hipLaunchKernelGGL((up_kernel<THi, C, N, KIND>), tg.grid, kBlock, 0, s, a);
  return hipGetLastError();
}

template <typename THi, int C>
static hipError_t up_c(int n, int kind, const UpLaunch& u, int batch, hipStream_t s) {
  if (kind == kBlend2) return n == 2 ? up_t<THi, C, 2, kBlend2>(u, batch, s) : hipErrorInvalidValue;
  if (kind == kBlend3) return n == 3 ? up_t<THi, C, 3, kBlend3>(u, batch, s) : hipErrorInvalidValue;
  switch (n) {
    case 2: return up_t<THi, C, 2, kBlendN>(u, batch, s);
    case 3: return up_t<THi, C, 3, kBlendN>(u, batch, s);
    case 4: return up_t<THi, C, 4, kBlendN>(u, batch, s);
    case 5: return up_t<THi, C, 5, kBlendN>(u, batch, s);
    case 6: return up_t<THi, C, 6, kBlendN>(u, batch, s);
    case 7: return up_t<THi, C, 7, kBlendN>(u, batch, s);
    case 8: return up_t<THi, C, 8, kBlendN>(u, batch, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_up(int c, int n, int kind, bool hi_u8, const UpLaunch& u, int batch, hipStream_t s) {
  if (c == 4) return hi_u8 ? up_c<uint8_t, 4>(n, kind, u, batch, s) : up_c<float, 4>(n, kind, u, batch, s);
  if (c == 3) return hi_u8 ? up_c<uint8_t, 3>(n, kind, u, batch, s) : up_c<float, 3>(n, kind, u, batch, s);
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// fused whole-pyramid kernels
// ---------------------------------------------------------------------------
template <typename T0, int C, int N>
static hipError_t down_fused_t(const void* const* g0, int w0, int h0, int n_out, const LevelDesc* lv,
                               const uint32_t* touch, int batch, hipStream_t s) {
  DownFusedArgs<T0, N> a;
  for (int i = 0; i < N; ++i) a.g0[i] = static_cast<const T0*>(g0[i]);
  fill_down_args(a, lv, w0, h0, n_out);
  a.touch = touch;
  const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
  if (g_recorder) return g_recorder->add(op_down<T0, C, N>(), grid, dim3(256), s, a, "down_fused");
  hipLaunchKernelGGL((down_fused_kernel<T0, C, N>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T0, int C>
static hipError_t down_fused_c(int n, const void* const* g0, int w0, int h0, int n_out, const LevelDesc* lv,
                               const uint32_t* touch,
                               int batch, hipStream_t s) {
  switch (n) {
    case 2: return down_fused_t<T0, C, 2>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 3: return down_fused_t<T0, C, 3>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 4: return down_fused_t<T0, C, 4>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 5: return down_fused_t<T0, C, 5>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 6: return down_fused_t<T0, C, 6>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 7: return down_fused_t<T0, C, 7>(g0, w0, h0, n_out, lv, touch, batch, s);
    case 8: return down_fused_t<T0, C, 8>(g0, w0, h0, n_out, lv, touch, batch, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_down_fused(int c, int n, bool in_u8, const void* const* g0, int w0, int h0, int n_out,
                             const LevelDesc* lv, const uint32_t* touch, int batch, hipStream_t s) {
  if (n_out < 1 || n_out > 5) return hipErrorInvalidValue;
  if (c == 4) return in_u8 ? down_fused_c<uint8_t, 4>(n, g0, w0, h0, n_out, lv, touch, batch, s)
                           : down_fused_c<float, 4>(n, g0, w0, h0, n_out, lv, touch, batch, s);
  if (c == 3) return in_u8 ? down_fused_c<uint8_t, 3>(n, g0, w0, h0, n_out, lv, touch, batch, s)
                           : down_fused_c<float, 3>(n, g0, w0, h0, n_out, lv, touch, batch, s);
  return hipErrorInvalidValue;
}

// Calls f(C, N, KIND) with std::integral_constant arguments for every
// supported (channels, images, blend kind).
template <typename F>
static hipError_t dispatch_cnk(int c, int n, int kind, F&& f) {
  using std::integral_constant;
  auto per_c = [&](auto C) -> hipError_t {
    if (kind == kBlend2)
      return n == 2 ? f(C, integral_constant<int, 2>{}, integral_constant<int, kBlend2>{}) : hipErrorInvalidValue;
    if (kind == kBlend3)
      return n == 3 ? f(C, integral_constant<int, 3>{}, integral_constant<int, kBlend3>{}) : hipErrorInvalidValue;
    if (kind != kBlendN) return hipErrorInvalidValue;
    const integral_constant<int, kBlendN> K{};
    switch (n) {
      case 2: return f(C, integral_constant<int, 2>{}, K);
      case 3: return f(C, integral_constant<int, 3>{}, K);
      case 4: return f(C, integral_constant<int, 4>{}, K);
      case 5: return f(C, integral_constant<int, 5>{}, K);
      case 6: return f(C, integral_constant<int, 6>{}, K);
      case 7: return f(C, integral_constant<int, 7>{}, K);
      case 8: return f(C, integral_constant<int, 8>{}, K);
      default: return hipErrorInvalidValue;
    }
  };
  if (c == 4) return per_c(integral_constant<int, 4>{});
  if (c == 3) return per_c(integral_constant<int, 3>{});
  return hipErrorInvalidValue;
}

// The burst tail of three or four images (3- and N-image blends).
template <int C, int N, int KIND>
static hipError_t up_tail_multi_t(const UpTailLaunch& u, int batch, hipStream_t s) {
  UpTailArgs<N> a;
  a.n = u.n;
  for (int j = 0; j < kMaxTail; ++j) {
    const int jj = j < u.n ? j : 0;
    for (int i = 0; i < N; ++i) {
      a.lv[j].g[i] = u.g[jj][i];
      a.lv[j].z[i] = u.z[jj][i];
    }
    a.lv[j].mask = u.mask[jj];
    a.lv[j].w = u.w[jj];
    a.lv[j].h = u.h[jj];
  }
  a.r_out = u.r_out;
  a.wx0 = u.wx0;
  a.wy0 = u.wy0;
  a.wx1 = u.wx1;
  a.wy1 = u.wy1;
  constexpr int T = kTailLdsTile;
  a.tiles_x = (u.wx1 - u.wx0 + T) / T;
  a.tiles_y = (u.wy1 - u.wy0 + T) / T;
  const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
  if (g_recorder) return g_recorder->fail("up_tail_lds");  // no batched form
  hipLaunchKernelGGL((up_tail_lds_kernel<C, N, KIND, T>), grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_up_tail(int c, int n, int kind, const UpTailLaunch& u, int batch, hipStream_t s) {
  if (n < 2 || n > kMaxTailImages || u.n < 2 || u.n > kMaxTail) return hipErrorInvalidValue;
  if (u.wx1 < u.wx0 || u.wy1 < u.wy0) return hipSuccess;
  if (n > 2) {
    if (u.n - 1 > kTailLds || (kind != kBlendN && kind != kBlend3) || (kind == kBlend3 && n != 3) ||
        (c != 3 && c != 4))
      return hipErrorInvalidValue;
    auto by_c = [&](auto Cv) -> hipError_t {
      constexpr int kC = decltype(Cv)::value;
      if (kind == kBlend3) return up_tail_multi_t<kC, 3, kBlend3>(u, batch, s);
      return n == 3 ? up_tail_multi_t<kC, 3, kBlendN>(u, batch, s) : up_tail_multi_t<kC, 4, kBlendN>(u, batch, s);
    };
    return c == 4 ? by_c(std::integral_constant<int, 4>{}) : by_c(std::integral_constant<int, 3>{});
  }
  UpTailArgs<2> a;
  a.n = u.n;
  for (int j = 0; j < kMaxTail; ++j) {
    const int jj = j < u.n ? j : 0;
    for (int i = 0; i < 2; ++i) {
      a.lv[j].g[i] = u.g[jj][i];
      a.lv[j].z[i] = u.z[jj][i];
    }
    a.lv[j].mask = u.mask[jj];
    a.lv[j].w = u.w[jj];
    a.lv[j].h = u.h[jj];
  }
  a.r_out = u.r_out;
  a.wx0 = u.wx0;
  a.wy0 = u.wy0;
  a.wx1 = u.wx1;
  a.wy1 = u.wy1;
  if (u.n - 1 <= kTailLds) {
    // every level's footprint in LDS, loaded in one burst
    constexpr int T = kTailLdsTile;
    a.tiles_x = (u.wx1 - u.wx0 + T) / T;
    a.tiles_y = (u.wy1 - u.wy0 + T) / T;
    const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
    if (g_recorder) {
      const int op =
          kind != kBlend2 ? kOpNone : (c == 4 ? op_up_tail_lds<4, 2, kBlend2>() : op_up_tail_lds<3, 2, kBlend2>());
      return g_recorder->add(op, grid, dim3(256), s, a, "up_tail_lds");
    }
    if (c == 4) {
      if (kind == kBlend2)
        hipLaunchKernelGGL((up_tail_lds_kernel<4, 2, kBlend2, T>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((up_tail_lds_kernel<4, 2, kBlendN, T>), grid, dim3(256), 0, s, a);
    } else if (c == 3) {
      if (kind == kBlend2)
        hipLaunchKernelGGL((up_tail_lds_kernel<3, 2, kBlend2, T>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((up_tail_lds_kernel<3, 2, kBlendN, T>), grid, dim3(256), 0, s, a);
    } else {
      return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  a.tiles_x = (u.wx1 - u.wx0 + kTailTile) / kTailTile;
  a.tiles_y = (u.wy1 - u.wy0 + kTailTile) / kTailTile;
  const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
  if (g_recorder) {
    const int op = kind != kBlend2 ? kOpNone : (c == 4 ? op_up_tail<4, 2, kBlend2>() : op_up_tail<3, 2, kBlend2>());
    return g_recorder->add(op, grid, dim3(256), s, a, "up_tail");
  }
  if (c == 4) {
    if (kind == kBlend2)
      hipLaunchKernelGGL((up_tail_kernel<4, 2, kBlend2>), grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((up_tail_kernel<4, 2, kBlendN>), grid, dim3(256), 0, s, a);
  } else if (c == 3) {
    if (kind == kBlend2)
      hipLaunchKernelGGL((up_tail_kernel<3, 2, kBlend2>), grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((up_tail_kernel<3, 2, kBlendN>), grid, dim3(256), 0, s, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Fills the level-0 arguments of a quad launch (window, tiles, regions).
template <typename T0, int N>
static void fill_quad_args(UpQuadArgs<T0, N>& a, const UpQuadLaunch& u) {
  for (int i = 0; i < N; ++i) {
    a.g0[i] = static_cast<const T0*>(u.g0[i]);
    a.g1[i] = u.g1[i];
  }
  a.mask0 = u.mask0;
  a.r1 = u.r1;
  a.w0 = u.w0;
  a.h0 = u.h0;
  a.w1 = u.w1;
  a.h1 = u.h1;
  a.f32_out = u.f32_out;
  a.u8_out = u.u8_out;
  a.u8_pitch = u.u8_pitch;
  a.u8_w = u.u8_w;
  a.u8_h = u.u8_h;
  a.ox = u.ox;
  a.oy = u.oy;
  a.win_x0 = u.win_x;
  a.win_y0 = u.win_y;
  a.win_x1 = u.win_x + u.win_w - 1;
  a.win_y1 = u.win_y + u.win_h - 1;
  a.qx0 = a.win_x0 >> 1;
  a.qy0 = a.win_y0 >> 1;
  a.nqx = (a.win_x1 >> 1) - a.qx0 + 1;
  a.nqy = (a.win_y1 >> 1) - a.qy0 + 1;
  a.tiles_x = (a.nqx + 63) / 64;
  a.tiles_y = (a.nqy + 3) / 4;
  for (int i = 0; i < N; ++i) a.z[i] = u.z[i];
  a.wn = u.wn;
  a.wn_plane = u.wn_plane;
}

hipError_t launch_up_quad(int c, int n, int kind, bool g0_u8, const UpQuadLaunch& u, int batch,
                            hipStream_t s) {
  if (u.win_w <= 0 || u.win_h <= 0) return hipSuccess;
  const int mode = u.mask0_mode;
  if (kind == kBlend2 ? (mode != kMask0F32 && mode != kMask0U8) : (mode != kMask0F32 && mode != kMask0Label))
    return hipErrorInvalidValue;
  return dispatch_cnk(c, n, kind, [&](auto C, auto N, auto K) -> hipError_t {
    constexpr int kC = decltype(C)::value, kN = decltype(N)::value, kK = decltype(K)::value;
    auto go = [&](auto T0v, auto M) -> hipError_t {
      using T0 = decltype(T0v);
      UpQuadArgs<T0, kN> a;
      fill_quad_args(a, u);
      const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
      if (g_recorder) {
        const int op = a.r1 ? op_up_quad<T0, kC, kN, kK, decltype(M)::value, true>()
                            : op_up_quad<T0, kC, kN, kK, decltype(M)::value, false>();
        return g_recorder->add(op, grid, dim3(64, 4, 1), s, a, "up_quad");
      }
      if (a.r1)
        hipLaunchKernelGGL((up_quad_kernel<T0, kC, kN, kK, decltype(M)::value, true>), grid, dim3(64, 4, 1), 0, s, a);
      else
        hipLaunchKernelGGL((up_quad_kernel<T0, kC, kN, kK, decltype(M)::value, false>), grid, dim3(64, 4, 1), 0, s,
                           a);
      return hipGetLastError();
    };
    using std::integral_constant;
    if constexpr (kK == kBlend2) {
      if (mode == kMask0U8)
        return g0_u8 ? go(uint8_t{}, integral_constant<int, kMask0U8>{}) : go(float{}, integral_constant<int, kMask0U8>{});
    } else {
      if (mode == kMask0Label)
        return g0_u8 ? go(uint8_t{}, integral_constant<int, kMask0Label>{})
                     : go(float{}, integral_constant<int, kMask0Label>{});
    }
    return g0_u8 ? go(uint8_t{}, integral_constant<int, kMask0F32>{}) : go(float{}, integral_constant<int, kMask0F32>{});
  });
}

hipError_t launch_up_l10(int c, bool g0_u8, const UpQuadLaunch& u, const float* mask1, const float* const* g2,
                         const float* r2, int w2, int h2, int batch, hipStream_t s) {
  if (u.win_w <= 0 || u.win_h <= 0) return hipSuccess;
  if (u.mask0_mode != kMask0F32 && u.mask0_mode != kMask0U8) return hipErrorInvalidValue;
  auto go = [&](auto T0v, auto Cv, auto M) -> hipError_t {
    using T0 = decltype(T0v);
    constexpr int kC = decltype(Cv)::value, kM = decltype(M)::value;
    UpL10Args<T0> A;
    UpQuadArgs<T0, 2>& a = A.l0;
    for (int i = 0; i < 2; ++i) {
      a.g0[i] = static_cast<const T0*>(u.g0[i]);
      a.g1[i] = u.g1[i];
      A.g2[i] = g2 ? g2[i] : nullptr;
    }
    a.mask0 = u.mask0;
    a.r1 = nullptr;
    a.w0 = u.w0;
    a.h0 = u.h0;
    a.w1 = u.w1;
    a.h1 = u.h1;
    a.f32_out = u.f32_out;
    a.u8_out = u.u8_out;
    a.u8_pitch = u.u8_pitch;
    a.u8_w = u.u8_w;
    a.u8_h = u.u8_h;
    a.ox = u.ox;
    a.oy = u.oy;
    a.win_x0 = u.win_x;
    a.win_y0 = u.win_y;
    a.win_x1 = u.win_x + u.win_w - 1;
    a.win_y1 = u.win_y + u.win_h - 1;
    a.qx0 = a.win_x0 >> 1;
    a.qy0 = a.win_y0 >> 1;
    a.nqx = (a.win_x1 >> 1) - a.qx0 + 1;
    a.nqy = (a.win_y1 >> 1) - a.qy0 + 1;
    a.tiles_x = (a.nqx + 63) / 64;
    a.tiles_y = (a.nqy + 3) / 4;
    for (int i = 0; i < 2; ++i) a.z[i] = u.z[i];
    a.wn = nullptr;
    a.wn_plane = 0;
    A.mask1 = mask1;
    A.r2 = r2;
    A.w2 = w2;
    A.h2 = h2;
    const dim3 grid((unsigned)(a.tiles_x * a.tiles_y), 1u, (unsigned)batch);
    if (g_recorder) {
      const int op = r2 ? op_up_l10<T0, kC, kM, true>() : op_up_l10<T0, kC, kM, false>();
      return g_recorder->add(op, grid, dim3(64, 4, 1), s, A, "up_l10");
    }
    if (r2)
      hipLaunchKernelGGL((up_l10_kernel<T0, kC, kM, true>), grid, dim3(64, 4, 1), 0, s, A);
    else
      hipLaunchKernelGGL((up_l10_kernel<T0, kC, kM, false>), grid, dim3(64, 4, 1), 0, s, A);
    return hipGetLastError();
  };
  using std::integral_constant;
  auto by_mode = [&](auto T0v, auto Cv) -> hipError_t {
    return u.mask0_mode == kMask0U8 ? go(T0v, Cv, integral_constant<int, kMask0U8>{})
                                    : go(T0v, Cv, integral_constant<int, kMask0F32>{});
  };
  auto by_c = [&](auto T0v) -> hipError_t {
    return c == 4 ? by_mode(T0v, integral_constant<int, 4>{}) : by_mode(T0v, integral_constant<int, 3>{});
  };
  if (c != 3 && c != 4) return hipErrorInvalidValue;
  return g0_u8 ? by_c(uint8_t{}) : by_c(float{});
}

bool l10n_window_fits(int n, const ImgRect* z0, const ImgRect* z1, int win_x, int win_y, int win_w, int win_h) {
  if (win_w <= 0 || win_h <= 0) return true;
  const int qx0 = win_x >> 1, qy0 = win_y >> 1;
  const int nqx = ((win_x + win_w - 1) >> 1) - qx0 + 1, nqy = ((win_y + win_h - 1) >> 1) - qy0 + 1;
  for (int ty = 0; ty < (nqy + 3) / 4; ++ty)
    for (int tx = 0; tx < (nqx + 63) / 64; ++tx)
      if (__builtin_popcount(l10n_tile_set(z0, z1, n, qx0 + tx * 64, qy0 + ty * 4)) > kQuadSlots) return false;
  return true;
}

hipError_t launch_up_l10n(int c, int n, const UpQuadLaunch& u, const float* wn1, size_t wn1_plane,
                          const ImgRect* z1, const float* const* g2, const float* r2, int w2, int h2, int batch,
                          hipStream_t s) {
  if (u.win_w <= 0 || u.win_h <= 0) return hipSuccess;
  if (u.mask0_mode != kMask0Label || n < 3 || n > kMaxImages || !wn1 || !r2 || (c != 3 && c != 4))
    return hipErrorInvalidValue;
  auto go = [&](auto Cv, auto Nv) -> hipError_t {
    constexpr int kC = decltype(Cv)::value, kN = decltype(Nv)::value;
    UpL10NArgs<kN> A;
    fill_quad_args(A.l0, u);
    A.l0.r1 = nullptr;
    A.l0.wn = nullptr;
    A.l0.wn_plane = 0;
    A.wn1 = wn1;
    A.wn1_plane = wn1_plane;
    for (int i = 0; i < kN; ++i) {
      A.z1[i] = z1[i];
      A.g2[i] = g2[i];
    }
    A.r2 = r2;
    A.w2 = w2;
    A.h2 = h2;
    const dim3 grid((unsigned)(A.l0.tiles_x * A.l0.tiles_y), 1u, (unsigned)batch);
    hipLaunchKernelGGL((up_l10n_kernel<kC, kN>), grid, dim3(64, 4, 1), 0, s, A);
    return hipGetLastError();
  };
  using std::integral_constant;
  auto by_n = [&](auto Cv) -> hipError_t {
    switch (n) {
      case 3: return go(Cv, integral_constant<int, 3>{});
      case 4: return go(Cv, integral_constant<int, 4>{});
      case 5: return go(Cv, integral_constant<int, 5>{});
      case 6: return go(Cv, integral_constant<int, 6>{});
      case 7: return go(Cv, integral_constant<int, 7>{});
      default: return go(Cv, integral_constant<int, 8>{});
    }
  };
  return c == 4 ? by_n(integral_constant<int, 4>{}) : by_n(integral_constant<int, 3>{});
}

// ---------------------------------------------------------------------------
// launch timing
// ---------------------------------------------------------------------------
LaunchTimer::~LaunchTimer() {
  for (auto& p : pending_) {
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
  if (cur_begin_) (void)hipEventDestroy(cur_begin_);
}

hipEvent_t LaunchTimer::get_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

void LaunchTimer::begin(hipStream_t s) {
  if (!enabled_) return;
  cur_begin_ = get_event();
  (void)hipEventRecord(cur_begin_, s);
}

void LaunchTimer::end(hipStream_t s, const char* name) {
  if (!enabled_ || !cur_begin_) return;
  hipEvent_t e = get_event();
  (void)hipEventRecord(e, s);
  pending_.push_back(Pending{cur_begin_, e, name});
  cur_begin_ = nullptr;
}

hipError_t LaunchTimer::collect() {
  for (auto& p : pending_) {
    HM_RETURN_IF_ERROR(hipEventSynchronize(p.b));
    float ms = 0.0f;
    HM_RETURN_IF_ERROR(hipEventElapsedTime(&ms, p.a, p.b));
    auto& t = totals_[p.name];
    t.first += ms;
    t.second += 1;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
  return hipSuccess;
}

std::string LaunchTimer::report() const {
  std::ostringstream os;
  for (const auto& kv : totals_) os << kv.first << " " << kv.second.first << " " << kv.second.second << "\n";
  return os.str();
}

// ---------------------------------------------------------------------------
// Pyramid
// ---------------------------------------------------------------------------
hipError_t Pyramid::init(int kind_, int n_, int channels, int width, int height, int levels_req, int batch_,
                         bool debug_, void* arena, size_t arena_bytes) {
  if ((kind_ != kBlend2 && kind_ != kBlend3 && kind_ != kBlendN) || n_ < 2 || n_ > kMaxImages ||
      (channels != 3 && channels != 4) || batch_ < 1)
    return hipErrorInvalidValue;
  if (arena && debug_) return hipErrorInvalidValue;
  if (kind_ == kBlend2 && n_ != 2) return hipErrorInvalidValue;
  if (kind_ == kBlend3 && n_ != 3) return hipErrorInvalidValue;
  kind = kind_;
  n = n_;
  c = channels;
  batch = batch_;
  debug = debug_;
  levels = pyramid_sizes(width, height, levels_req, kind != kBlend2, W, H);
  if (levels < 1) return hipErrorInvalidValue;
  allocation_size = 0;
  for (int l = 0; l < kMaxLevels; ++l) {
    wn[l].reset();
    touch[l].reset();
  }
  // Arena layout: every level's buffers in the order below, each at a
  // 256-byte boundary.  First the size check, then the views.
  const auto round_up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  arena_need = 0;
  for (int l = 1; l < levels; ++l) {
    const size_t px = (size_t)W[l] * H[l];
    const size_t img = round_up(px * c * batch * sizeof(float));
    arena_need += (n + 1) * img + round_up(px * mask_channels() * sizeof(float));
    if (kind == kBlendN && n > 2) arena_need += round_up(px * n * sizeof(float));
  }
  if (arena && arena_need > arena_bytes) return hipErrorOutOfMemory;
  char* cursor = static_cast<char*>(arena);
  const auto take = [&](DevBuf& b, size_t bytes) -> hipError_t {
    if (!arena) return b.alloc(bytes);
    b.view(cursor, bytes);
    cursor += round_up(bytes);
    return hipSuccess;
  };
  for (int l = 0; l < levels; ++l) {
    const size_t px = (size_t)W[l] * H[l];
    const size_t img = px * c * batch * sizeof(float);
    if (l > 0) {
      for (int i = 0; i < n; ++i) HM_RETURN_IF_ERROR(take(g[i][l], img));
      HM_RETURN_IF_ERROR(take(mask[l], px * mask_channels() * sizeof(float)));
      HM_RETURN_IF_ERROR(take(r[l], img));
      allocation_size += n * img + img + px * mask_channels() * sizeof(float);
      if (kind == kBlendN && n > 2 && !debug) {  // the slot path of three or more images
        HM_RETURN_IF_ERROR(take(wn[l], px * n * sizeof(float)));
        allocation_size += px * n * sizeof(float);
      }
    }
    if (debug) {
      for (int i = 0; i < n; ++i) HM_RETURN_IF_ERROR(lap[i][l].alloc(img));
      HM_RETURN_IF_ERROR(blend[l].alloc(img));
      allocation_size += (n + 1) * img;
    }
  }
  if (debug) {
    HM_RETURN_IF_ERROR(r0_debug.alloc((size_t)W[0] * H[0] * c * batch * sizeof(float)));
    allocation_size += r0_debug.bytes;
  }
  // dense until set_supports
  for (int l = 0; l < levels; ++l)
    for (int i = 0; i < kMaxImages; ++i) S[l][i] = Z[l][i] = ImgRect{0, 0, W[l] - 1, H[l] - 1};
  build_level_table();
  return hipSuccess;
}

void Pyramid::build_level_table() {
  for (int l = 0; l < kMaxLevels; ++l) level_table[l] = LevelDesc{};
  for (int l = 1; l < levels; ++l) {
    LevelDesc& d = level_table[l - 1];
    for (int i = 0; i < kMaxImages; ++i) {
      d.g[i] = i < n ? g[i][l].as<float>() : nullptr;
      d.g_out[i] = i < n ? g[i][l].as<float>() : nullptr;
      d.s[i] = S[l][i];
      d.z[i] = Z[l][i];
    }
    d.mask = mask[l].as<float>();
    d.w = W[l];
    d.h = H[l];
  }
}

hipError_t Pyramid::set_supports(const ImgRect* s0) {
  l10n_key[0] = -1;  // the regions change: find the fused up launch's fit again
  // The debug path keeps every level dense, and so do one or two images (the
  // 2-image quad path reads every image everywhere).
  if (debug || n <= 2) return hipSuccess;
  for (int i = 0; i < n; ++i) {
    ImgRect s = s0[i];
    for (int l = 0; l < levels; ++l) {
      if (l > 0 && s.x1 >= s.x0 && s.y1 >= s.y0) s = ImgRect{s.x0 >> 1, s.y0 >> 1, s.x1 >> 1, s.y1 >> 1};
      // clip to the level, then grow by 2 for the valid region
      ImgRect c{std::max(s.x0, 0), std::max(s.y0, 0), std::min(s.x1, W[l] - 1), std::min(s.y1, H[l] - 1)};
      const bool empty = c.x1 < c.x0 || c.y1 < c.y0;
      S[l][i] = empty ? ImgRect{0, 0, -1, -1} : c;
      Z[l][i] = empty ? ImgRect{0, 0, -1, -1}
                      : ImgRect{std::max(c.x0 - 2, 0), std::max(c.y0 - 2, 0), std::min(c.x1 + 2, W[l] - 1),
                                std::min(c.y1 + 2, H[l] - 1)};
      s = c;
      if (empty) s = ImgRect{0, 0, -1, -1};
    }
  }
  build_level_table();
  // Per down launch (levels from+1 .. from+n_out out of one launch), the
  // images each tile meets: the same test the kernel makes without a table.
  for (int from = 0; from + 1 < levels;) {
    const int n_out = std::min(5, levels - 1 - from);
    const int tiles_x = (W[from] + kDownTile - 1) / kDownTile, tiles_y = (H[from] + kDownTile - 1) / kDownTile;
    std::vector<uint32_t> tab((size_t)tiles_x * tiles_y, 0u);
    for (int ty = 0; ty < tiles_y; ++ty)
      for (int tx = 0; tx < tiles_x; ++tx) {
        uint32_t act = 0;
        for (int i = 0; i < n; ++i)
          for (int k = 0; k < n_out; ++k)
            if (tile_meets(Z[from + 1 + k][i], tx, ty, 16 >> k)) act |= 1u << i;
        tab[(size_t)ty * tiles_x + tx] = act;
      }
    HM_RETURN_IF_ERROR(touch[from].alloc(tab.size() * sizeof(uint32_t)));
    HM_RETURN_IF_ERROR(hipMemcpy(touch[from].p, tab.data(), touch[from].bytes, hipMemcpyHostToDevice));
    allocation_size += touch[from].bytes;
    from += n_out;
  }
  return hipSuccess;
}

hipError_t Pyramid::build_mask_pyramid(const float* mask0_, hipStream_t s) {
  mask0 = mask0_;
  const float* prev = mask0;
  for (int l = 0; l + 1 < levels; ++l) {
    HM_RETURN_IF_ERROR(
        launch_mask_down(prev, W[l], H[l], mask[l + 1].as<float>(), W[l + 1], H[l + 1], mask_channels(), s));
    prev = mask[l + 1].as<float>();
    if (wn[l + 1].p)
      HM_RETURN_IF_ERROR(launch_mask_planar(prev, W[l + 1], H[l + 1], wn[l + 1].as<float>(), n, s));
  }
  return hipSuccess;
}

hipError_t Pyramid::run_down(const void* const* g0, bool g0_u8, int from, hipStream_t s, LaunchTimer* timer) {
  // Up to five levels per launch from aligned tiles.
  const LevelDesc* table = level_table;
  while (from + 1 < levels) {
    const int n_out = std::min(5, levels - 1 - from);
    const void* in[kMaxImages];
    for (int i = 0; i < n; ++i) in[i] = from == 0 ? g0[i] : g[i][from].p;
    if (timer) timer->begin(s);
    HM_RETURN_IF_ERROR(launch_down_fused(c, n, from == 0 && g0_u8, in, W[from], H[from], n_out, table + from,
                                         touch[from].as<uint32_t>(), batch, s));
    if (timer) timer->end(s, "down_fused");
    from += n_out;
  }
  return hipSuccess;
}

hipError_t Pyramid::run_up(const void* const* g0, bool g0_u8, const Level0Out& out, hipStream_t s,
                           LaunchTimer* timer) {
  const int wx = out.win_w >= 0 ? out.win_x : 0, wy = out.win_w >= 0 ? out.win_y : 0;
  const int ww = out.win_w >= 0 ? out.win_w : W[0], wh = out.win_w >= 0 ? out.win_h : H[0];
  if (ww <= 0 || wh <= 0) return hipSuccess;
  // Footprint of the level-0 window at every level: the pixels whose R the
  // next finer level's taps read.
  int fx0[kMaxLevels], fx1[kMaxLevels], fy0[kMaxLevels], fy1[kMaxLevels];
  fx0[0] = wx;
  fy0[0] = wy;
  fx1[0] = wx + ww - 1;
  fy1[0] = wy + wh - 1;
  for (int k = 1; k < levels; ++k) {
    fx0[k] = fx0[k - 1] >> 1;
    fy0[k] = fy0[k - 1] >> 1;
    fx1[k] = std::min((fx1[k - 1] >> 1) + 1, W[k] - 1);
    fy1[k] = std::min((fy1[k - 1] >> 1) + 1, H[k] - 1);
  }
  // Two images: levels tail_base .. L-1 in one tile-fused launch, giving R at
  // tail_base on the footprint the finer levels read.
  // Three or four images (sparse, the 3- and N-image blends): the same
  // tail in its burst form, loading each image inside its valid region only,
  // when no batched pass is being recorded.
  int top = levels - 1;  // first level run in quads
  const int tb = tail_base;
  const bool tail_multi = n > 2 && n <= kMaxTailImages && (kind == kBlendN || kind == kBlend3) &&
                          levels - 1 - tb <= kTailLds && !g_recorder;
  if ((n == 2 || tail_multi) && tb > 0 && levels - 1 > tb && levels - tb <= kMaxTail) {
    UpTailLaunch t;
    t.n = levels - tb;
    for (int j = 0; j < t.n; ++j) {
      const int l = tb + j;
      for (int i = 0; i < n; ++i) {
        t.g[j][i] = g[i][l].as<float>();
        t.z[j][i] = Z[l][i];
      }
      t.mask[j] = mask[l].as<float>();
      t.w[j] = W[l];
      t.h[j] = H[l];
    }
    t.r_out = r[tb].as<float>();
    t.wx0 = fx0[tb];
    t.wy0 = fy0[tb];
    t.wx1 = fx1[tb];
    t.wy1 = fy1[tb];
    if (timer) timer->begin(s);
    HM_RETURN_IF_ERROR(launch_up_tail(c, n, kind, t, batch, s));
    if (timer) timer->end(s, "up_tail");
    top = tb - 1;
  }
  // The other levels one launch each, coarsest first, one thread per 2x2
  // quad; level 0 goes straight to the output.  The 2-image blend: levels 1
  // and 0 in one launch (fuse_up) when level 1 is not the tail's.
  const bool l10 = kind == kBlend2 && fuse_up && top >= 1 && levels >= 2;
  // Three or more images with uint8 labels at level 0: levels 1 and 0 in one
  // launch too (fuse_up), when no tile has more images than slots (checked
  // once per window) and no batched pass is being recorded.
  bool l10n = false;
  if (kind == kBlendN && n > 2 && fuse_up && g0_u8 && top >= 1 && levels >= 3 && mask0_fast &&
      mask0_mode == kMask0Label && wn[1].p && !g_recorder) {
    const int key[4] = {fx0[0], fy0[0], fx1[0], fy1[0]};
    if (!std::equal(key, key + 4, l10n_key)) {
      l10n_fits = l10n_window_fits(n, Z[0], Z[1], fx0[0], fy0[0], fx1[0] - fx0[0] + 1, fy1[0] - fy0[0] + 1);
      std::copy(key, key + 4, l10n_key);
    }
    l10n = l10n_fits;
  }
  for (int k = top; k >= 0; --k) {
    if ((l10 || l10n) && k == 1) continue;  // computed inside the level-0 launch
    const bool has_next = k + 1 < levels;
    UpQuadLaunch u;
    for (int i = 0; i < n; ++i) {
      u.g0[i] = k == 0 ? g0[i] : g[i][k].p;
      u.g1[i] = has_next ? g[i][k + 1].as<float>() : nullptr;
    }
    if (k == 0) {
      u.mask0 = mask0_fast ? mask0_fast : mask0;
      u.mask0_mode = mask0_fast ? mask0_mode : kMask0F32;
    } else {
      u.mask0 = mask[k].p;
      u.mask0_mode = kMask0F32;
      u.wn = wn[k].as<float>();
      u.wn_plane = (size_t)W[k] * H[k];
    }
    u.r1 = has_next ? r[k + 1].as<float>() : nullptr;
    u.w0 = W[k];
    u.h0 = H[k];
    u.w1 = has_next ? W[k + 1] : 0;
    u.h1 = has_next ? H[k + 1] : 0;
    if (k == 0) {
      u.f32_out = out.f32;
      u.u8_out = out.u8;
      u.u8_pitch = out.u8_pitch;
      u.u8_w = out.u8_w;
      u.u8_h = out.u8_h;
      u.ox = out.ox;
      u.oy = out.oy;
    } else {
      u.f32_out = r[k].as<float>();
    }
    u.win_x = fx0[k];
    u.win_y = fy0[k];
    u.win_w = fx1[k] - fx0[k] + 1;
    u.win_h = fy1[k] - fy0[k] + 1;
    for (int i = 0; i < n; ++i) u.z[i] = Z[k][i];
    if (timer) timer->begin(s);
    if (l10n && k == 0) {
      const float* g2[kMaxImages];
      for (int i = 0; i < n; ++i) g2[i] = g[i][2].as<float>();
      HM_RETURN_IF_ERROR(launch_up_l10n(c, n, u, wn[1].as<float>(), (size_t)W[1] * H[1], Z[1], g2,
                                        r[2].as<float>(), W[2], H[2], batch, s));
      if (timer) timer->end(s, "up_l10");
      continue;
    }
    if (l10 && k == 0) {
      const float* g2[2] = {levels > 2 ? g[0][2].as<float>() : nullptr, levels > 2 ? g[1][2].as<float>() : nullptr};
      HM_RETURN_IF_ERROR(launch_up_l10(c, g0_u8, u, mask[1].as<float>(), g2, levels > 2 ? r[2].as<float>() : nullptr,
                                       levels > 2 ? W[2] : 0, levels > 2 ? H[2] : 0, batch, s));
      if (timer) timer->end(s, "up_l10");
      continue;
    }
    HM_RETURN_IF_ERROR(launch_up_quad(c, n, kind, k == 0 && g0_u8, u, batch, s));
    if (timer) timer->end(s, k == 0 ? "up_level0" : "up_quad");
  }
  return hipSuccess;
}

hipError_t Pyramid::run(const void* const* g0, bool g0_u8, const Level0Out& out, hipStream_t s, LaunchTimer* timer) {
  if (!debug) {
    HM_RETURN_IF_ERROR(run_down(g0, g0_u8, 0, s, timer));
    return run_up(g0, g0_u8, out, s, timer);
  }
  // Debug: per-level launches that also keep the Laplacian and blended levels.
  // 1. Down pass: every image's level k -> k+1 in one launch per level.
  for (int l = 0; l + 1 < levels; ++l) {
    const void* in[kMaxImages];
    float* o[kMaxImages];
    for (int i = 0; i < n; ++i) {
      in[i] = l == 0 ? g0[i] : g[i][l].p;
      o[i] = g[i][l + 1].as<float>();
    }
    if (timer) timer->begin(s);
    HM_RETURN_IF_ERROR(launch_down(c, n, l == 0 && g0_u8, in, o, W[l], H[l], W[l + 1], H[l + 1], batch, s));
    if (timer) timer->end(s, l == 0 ? "down_l0" : "down");
  }
  // 2. Up pass from the coarsest level: Laplacian + blend + reconstruct fused.
  for (int l = levels - 1; l >= 0; --l) {
    UpLaunch u;
    const bool coarsest = l == levels - 1;
    for (int i = 0; i < n; ++i) {
      u.g_hi[i] = l == 0 ? g0[i] : g[i][l].p;
      u.g_lo[i] = coarsest ? nullptr : g[i][l + 1].as<float>();
      u.lap_out[i] = debug ? lap[i][l].as<float>() : nullptr;
    }
    u.mask = l == 0 ? mask0 : mask[l].as<float>();
    u.r_lo = coarsest ? nullptr : r[l + 1].as<float>();
    u.blend_out = debug ? blend[l].as<float>() : nullptr;
    u.hi_w = W[l];
    u.hi_h = H[l];
    u.lo_w = coarsest ? 0 : W[l + 1];
    u.lo_h = coarsest ? 0 : H[l + 1];
    u.win_x = 0;
    u.win_y = 0;
    u.win_w = W[l];
    u.win_h = H[l];
    if (l > 0) {
      u.r_out = r[l].as<float>();
    } else {
      if (debug) {
        // Full level-0 evaluation into the debug buffers only; the real
        // output is written by the (possibly windowed) launch below.
        u.r_out = r0_debug.as<float>();
        HM_RETURN_IF_ERROR(launch_up(c, n, kind, g0_u8, u, batch, s));
        for (int i = 0; i < n; ++i) u.lap_out[i] = nullptr;
        u.blend_out = nullptr;
      }
      u.r_out = out.f32;
      u.u8_out = out.u8;
      u.u8_pitch = out.u8_pitch;
      u.u8_w = out.u8_w;
      u.u8_h = out.u8_h;
      u.ox = out.ox;
      u.oy = out.oy;
      if (out.win_w >= 0) {
        u.win_x = out.win_x;
        u.win_y = out.win_y;
        u.win_w = out.win_w;
        u.win_h = out.win_h;
      }
    }
    if (timer) timer->begin(s);
    HM_RETURN_IF_ERROR(launch_up(c, n, kind, l == 0 && g0_u8, u, batch, s));
    if (timer) timer->end(s, l == 0 ? "up_l0" : "up");
  }
  return hipSuccess;
}

}  // namespace hm

// ===========================================================================
// Kernel-tier C-ABI: Laplacian blend contexts
// ===========================================================================
struct hmBlendContext {
  hm::Pyramid pyr;
  int width = 0, height = 0;
  hm::LaunchTimer timer;
};

extern "C" int hmBlendContextCreate(int kind, int n_images, int width, int height, int num_levels, int batch_size,
                                    int channels, int debug, hmBlendContext** out) {
  if (!out) return hipErrorInvalidValue;
  *out = nullptr;
  if (kind != HM_BLEND_TWO && kind != HM_BLEND_THREE && kind != HM_BLEND_N) return hipErrorInvalidValue;
  if (width < 1 || height < 1 || num_levels < 1) return hipErrorInvalidValue;
  auto* ctx = new hmBlendContext();
  const int k = kind == HM_BLEND_TWO ? hmdev::kBlend2 : (kind == HM_BLEND_THREE ? hmdev::kBlend3 : hmdev::kBlendN);
  hipError_t e = ctx->pyr.init(k, n_images, channels, width, height, num_levels, batch_size, debug != 0);
  if (e != hipSuccess) {
    delete ctx;
    return e;
  }
  ctx->width = width;
  ctx->height = height;
  *out = ctx;
  return hipSuccess;
}

extern "C" int hmBlendContextLevels(const hmBlendContext* ctx, int* num_levels, int* widths, int* heights) {
  if (!ctx || !num_levels) return hipErrorInvalidValue;
  *num_levels = ctx->pyr.levels;
  for (int l = 0; l < ctx->pyr.levels; ++l) {
    if (widths) widths[l] = ctx->pyr.W[l];
    if (heights) heights[l] = ctx->pyr.H[l];
  }
  return hipSuccess;
}

extern "C" int hmBlendRun(hmBlendContext* ctx, const float* const* d_images, const float* d_mask, float* d_output,
                          hipStream_t stream) {
  if (!ctx || !d_images || !d_mask || !d_output) return hipErrorInvalidValue;
  const int n = ctx->pyr.n;
  const void* g0[hmdev::kMaxImages];
  for (int i = 0; i < n; ++i) {
    if (!d_images[i]) return hipErrorInvalidValue;
    if (ctx->pyr.c == 4 && (reinterpret_cast<uintptr_t>(d_images[i]) & 15u)) return hipErrorInvalidValue;
    g0[i] = d_images[i];
  }
  if (ctx->pyr.c == 4 && (reinterpret_cast<uintptr_t>(d_output) & 15u)) return hipErrorInvalidValue;
  // The reference recomputes the mask pyramid on every call (cudaBlend.cu:912-923).
  HM_RETURN_IF_ERROR(ctx->pyr.build_mask_pyramid(d_mask, stream));
  hm::Level0Out out;
  out.f32 = d_output;
  return ctx->pyr.run(g0, false, out, stream, ctx->timer.enabled() ? &ctx->timer : nullptr);
}

extern "C" int hmBlendContextLevelPtr(const hmBlendContext* ctx, int what, int image, int level, const float** ptr) {
  if (!ctx || !ptr || level < 0 || level >= ctx->pyr.levels) return hipErrorInvalidValue;
  const hm::Pyramid& p = ctx->pyr;
  if (image < 0 || image >= p.n) image = 0;
  *ptr = nullptr;
  switch (what) {
    case HM_LEVEL_GAUSS: *ptr = level ? p.g[image][level].as<float>() : nullptr; break;
    case HM_LEVEL_MASK: *ptr = level ? p.mask[level].as<float>() : p.mask0; break;
    case HM_LEVEL_LAPLACIAN: *ptr = p.debug ? p.lap[image][level].as<float>() : nullptr; break;
    case HM_LEVEL_BLEND: *ptr = p.debug ? p.blend[level].as<float>() : nullptr; break;
    case HM_LEVEL_RECONSTRUCT: *ptr = level ? p.r[level].as<float>() : (p.debug ? p.r0_debug.as<float>() : nullptr); break;
    default: return hipErrorInvalidValue;
  }
  return hipSuccess;
}

extern "C" void hmBlendContextDestroy(hmBlendContext* ctx) { delete ctx; }

// ===========================================================================
// Allocation-free blend: the pyramid levels live in `scratch`, a file-scope
// device array (one copy per GPU, held until the process ends).
// ===========================================================================
static __device__ __align__(256) unsigned char scratch[HM_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch[kMaxDevices];  // device address of `scratch`, per device
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device, looked up once per
// device.  The array's name is a host-side stand-in: only this address may go
// to kernels or copies.
hipError_t scratch_ptr(void** out) {
  int dev = -1;
  HM_RETURN_IF_ERROR(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  void* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      HM_RETURN_IF_ERROR(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (!p) return hipErrorInvalidSymbol;
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return hipSuccess;
}
}  // namespace

extern "C" int hmBlendRunNoAlloc(int kind, int n_images, int width, int height, int num_levels, int batch_size,
                                 int channels, const float* const* d_images, const float* d_mask, float* d_output,
                                 hipStream_t stream) {
  if (kind != HM_BLEND_TWO && kind != HM_BLEND_THREE && kind != HM_BLEND_N) return hipErrorInvalidValue;
  if (width < 1 || height < 1 || num_levels < 1) return hipErrorInvalidValue;
  if (!d_images || !d_mask || !d_output) return hipErrorInvalidValue;
  void* arena = nullptr;
  HM_RETURN_IF_ERROR(scratch_ptr(&arena));
  // Host state only (fixed-size arrays); the device buffers are views of the scratch.
  hm::Pyramid pyr;
  const int k = kind == HM_BLEND_TWO ? hmdev::kBlend2 : (kind == HM_BLEND_THREE ? hmdev::kBlend3 : hmdev::kBlendN);
  HM_RETURN_IF_ERROR(pyr.init(k, n_images, channels, width, height, num_levels, batch_size, false, arena,
                              HM_NOALLOC_SCRATCH_BYTES));
  const void* g0[hmdev::kMaxImages];
  for (int i = 0; i < pyr.n; ++i) {
    if (!d_images[i]) return hipErrorInvalidValue;
    if (pyr.c == 4 && (reinterpret_cast<uintptr_t>(d_images[i]) & 15u)) return hipErrorInvalidValue;
    g0[i] = d_images[i];
  }
  if (pyr.c == 4 && (reinterpret_cast<uintptr_t>(d_output) & 15u)) return hipErrorInvalidValue;
  // Every level is written in full before it is read (mask pyramid, down pass,
  // up pass), so nothing in the scratch needs clearing.
  HM_RETURN_IF_ERROR(pyr.build_mask_pyramid(d_mask, stream));
  hm::Level0Out out;
  out.f32 = d_output;
  return pyr.run(g0, false, out, stream, nullptr);
}
