// hm_pyramid_kernels.hpp -- Gaussian/Laplacian pyramid kernels for gfx950.
//
// Down pass: one thread per coarse pixel, all N images in one launch
//   (FusedBatchedDownsampleKernel, src/cuda/cudaBlend.cu:46-144; the N- and
//   3-image forms cudaBlendN.cu:53-113, cudaBlend3.cu:100-215 are the same
//   arithmetic).
// Up pass: one fused kernel per level computes, for every pixel of level k,
//   Laplacian(k) of every image (cudaBlend.cu:356-456), the seam-weighted
//   blend (2-image :520-576, N-image cudaBlendN.cu:199-288, 3-image
//   cudaBlend3.cu:376-531) and the reconstruction R_k = up(R_k+1) + B_k
//   (cudaBlend.cu:581-672), so the Laplacian and blend pyramids never touch
//   HBM.  At level 0 it can store the truncated uint8 result straight into the
//   caller's canvas (the copy-back of cudaPano.inl:401-411).
#pragma once

#include "hm_device.hpp"

namespace hmdev {

constexpr int kMaxImages = 8;
constexpr int kMaxLevels = 16;

enum BlendKind { kBlendN = 0, kBlend2 = 2, kBlend3 = 3 };

// ---------------------------------------------------------------------------
// Down pass
// ---------------------------------------------------------------------------
template <typename TIn, int N>
struct DownArgs {
  const TIn* in[N];
  float* out[N];
  int in_w, in_h, out_w, out_h;
  int tiles_x, tiles_y;
};

// x / d for a tap count d in {0, 1, 2, 3, 4}, y = RN(1 / d) (0 for d = 0,
// where x = +0): q0 = x * y, the residual r = x - d * q0 (exact, one fused
// multiply-add) and q0 + r * y rounded once.  For d = 3 this is Markstein's
// correction step, exact because y is the correctly rounded reciprocal; for
// d = 1, 2, 4 q0 is the quotient itself (r = 0) except where x / d is
// subnormal.  Checked on every finite float x but -0 (which a sum that
// starts at +0 never is) against IEEE division, on the CPU: the same bits as
// x / d, for about a third of the division sequence's instructions.
__device__ __forceinline__ float div_count(float x, float d, float y) {
  const float q0 = x * y;
  const float r = __builtin_fmaf(-d, q0, x);
  return __builtin_fmaf(r, y, q0);
}

// One coarse pixel from its 2x2 block, taps in the reference order (0,0),
// (1,0), (0,1), (1,1): RGB mean of the in-range taps whose alpha != 0 (0 if
// none), alpha = max over the in-range taps (cudaBlend.cu:85-134).  Written
// without branches: a tap that does not count adds +0, which leaves a sum
// that starts at +0 unchanged (under round to nearest it never becomes -0),
// so the sums are the conditional sums' bits.  Taps out of range are never
// read (p[k] need not hold a value there).
template <int C>
__device__ __forceinline__ float4 down_px(const float4* p, const bool* in_range) {
  float sx = 0.0f, sy = 0.0f, sz = 0.0f, alpha = 0.0f;
  int count = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const bool in = in_range[k];
    const float4 q = in ? p[k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool keep = in;
    if constexpr (C == 4) {
      alpha = (in && alpha < q.w) ? q.w : alpha;
      keep = in && q.w != 0.0f;
    }
    sx += keep ? q.x : 0.0f;
    sy += keep ? q.y : 0.0f;
    sz += keep ? q.z : 0.0f;
    count += keep ? 1 : 0;
  }
  const float y = count == 4 ? 0.25f : count == 3 ? 0x1.555556p-2f : count == 2 ? 0.5f : count == 1 ? 1.0f : 0.0f;
  const float d = (float)count;
  return make_float4(div_count(sx, d, y), div_count(sy, d, y), div_count(sz, d, y), alpha);
}

template <typename TIn, int C, int N>
__global__ void __launch_bounds__(256) down_kernel(DownArgs<TIn, N> a) {
  const TileXY p0 = tile_pixel(a.tiles_x, a.tiles_y);
  const int x = p0.x, y = p0.y;
  const int b = blockIdx.z;
  if (x >= a.out_w || y >= a.out_h) return;
  const int ix = 2 * x, iy = 2 * y;
  const bool has_x1 = ix + 1 < a.in_w, has_y1 = iy + 1 < a.in_h;
  const size_t r0 = ((size_t)b * a.in_h + iy) * a.in_w;
  const size_t r1 = r0 + a.in_w;
  const size_t o = ((size_t)b * a.out_h + y) * a.out_w + x;
  const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
#pragma unroll
  for (int i = 0; i < N; ++i) {
    float4 p[4];
    p[0] = ld_px<C>(a.in[i], r0 + ix);
    if (has_x1) p[1] = ld_px<C>(a.in[i], r0 + ix + 1);
    if (has_y1) p[2] = ld_px<C>(a.in[i], r1 + ix);
    if (has_x1 && has_y1) p[3] = ld_px<C>(a.in[i], r1 + ix + 1);
    st_px<C>(a.out[i], o, down_px<C>(p, in_range));
  }
}

// ---------------------------------------------------------------------------
// Blend of one pixel's Laplacians
// ---------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ float4 blend2_px(const float4& l1, const float4& l2, float m) {
  const float mm1 = 1.0f - m;
  if constexpr (C == 4) {
    if (l1.w == 0.0f) return l2;
    if (l2.w == 0.0f) return l1;
    return make_float4(m * l1.x + mm1 * l2.x, m * l1.y + mm1 * l2.y, m * l1.z + mm1 * l2.z, l2.w);
  } else {
    return make_float4(m * l1.x + mm1 * l2.x, m * l1.y + mm1 * l2.y, m * l1.z + mm1 * l2.z, 0.0f);
  }
}

// The N-image blend after its first normalisation: w[i] = mask_i / sum of
// the pixel's mask values (the mask values themselves when that sum is 0).
// w is modified.  An image whose Laplacian is zero at the pixel (alpha 0, or
// RGB zeros) adds nothing here beyond what the mask sum already took from it:
// its weight is dropped (RGBA) or multiplies zeros (RGB), and the running sums
// start at +0 and never become -0 (rounding to nearest), so adding a +-0 term
// changes nothing.  A caller may therefore pass only the images that can be
// non-zero at the pixel, in index order.
template <int C, int N>
__device__ __forceinline__ float4 blendN_weighted(const float4* L, float* w) {
  if constexpr (C == 4) {
    float sum_w2 = 0.0f, max_alpha = 0.0f;
    int max_idx = -1;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const float a = L[i].w;
      if (a > max_alpha) {
        max_alpha = a;
        max_idx = i;
      }
      if (a == 0.0f) w[i] = 0.0f;
      sum_w2 += w[i];
    }
    if (sum_w2 > 0.0f) {
#pragma unroll
      for (int i = 0; i < N; ++i) w[i] /= sum_w2;
    } else {
      float4 r = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (max_idx >= 0 && max_alpha != 0.0f) {
#pragma unroll
        for (int i = 0; i < N; ++i)
          if (i == max_idx) r = L[i];
      }
      return r;
    }
  }
  float ax = 0.0f, ay = 0.0f, az = 0.0f, aw = 0.0f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ax += w[i] * L[i].x;
    ay += w[i] * L[i].y;
    az += w[i] * L[i].z;
  }
  if constexpr (C == 4) {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (w[i] == 0.0f) continue;
      aw = aw < L[i].w ? L[i].w : aw;
    }
  }
  return make_float4(ax, ay, az, aw);
}

// First normalisation of the N-image blend: the mask values over their sum
// when that is positive (cudaBlendN.cu:216-223).
template <int N>
__device__ __forceinline__ void blendN_normalize(const float* mp, float* w) {
  float sum_w = 0.0f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    w[i] = mp[i];
    sum_w += w[i];
  }
  if (sum_w > 0.0f) {
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] /= sum_w;
  }
}

template <int C, int N>
__device__ __forceinline__ float4 blendN_px(const float4* L, const float* mp) {
  float w[N];
  blendN_normalize<N>(mp, w);
  return blendN_weighted<C, N>(L, w);
}

template <int C>
__device__ __forceinline__ float4 blend3_px(const float4* L, const float* mp) {
  float m1 = mp[0], m2 = mp[1], m3 = mp[2];
  if constexpr (C == 4) {
    const float a1 = L[0].w, a2 = L[1].w, a3 = L[2].w;
    if (a1 == 0.0f) m1 = 0.0f;
    if (a2 == 0.0f) m2 = 0.0f;
    if (a3 == 0.0f) m3 = 0.0f;
    const float sum = m1 + m2 + m3;
    if (sum > 0.0f) {
      m1 /= sum;
      m2 /= sum;
      m3 /= sum;
    } else {
      float amax = 0.0f;
      int src = -1;
      if (a1 > amax) { amax = a1; src = 0; }
      if (a2 > amax) { amax = a2; src = 1; }
      if (a3 > amax) { amax = a3; src = 2; }
      if (src < 0) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      return src == 0 ? L[0] : (src == 1 ? L[1] : L[2]);
    }
  }
  float4 r;
  r.x = m1 * L[0].x + m2 * L[1].x + m3 * L[2].x;
  r.y = m1 * L[0].y + m2 * L[1].y + m3 * L[2].y;
  r.z = m1 * L[0].z + m2 * L[1].z + m3 * L[2].z;
  r.w = 0.0f;
  if constexpr (C == 4) {
    float a_out = 0.0f;
    if (m1 > 0.0f) a_out = a_out < L[0].w ? L[0].w : a_out;
    if (m2 > 0.0f) a_out = a_out < L[1].w ? L[1].w : a_out;
    if (m3 > 0.0f) a_out = a_out < L[2].w ? L[2].w : a_out;
    r.w = a_out;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Fused up pass
// ---------------------------------------------------------------------------
template <typename THi, int N>
struct UpArgs {
  const THi* g_hi[N];     // level-k Gaussians (uint8 at level 0 of a pano)
  const float* g_lo[N];   // level-(k+1) Gaussians; [0] == nullptr at the coarsest level
  const float* mask;      // level-k mask, M interleaved channels, unbatched
  const float* r_lo;      // level-(k+1) reconstruction; nullptr at the coarsest level
  float* lap_out[N];      // optional (debug dump of the Laplacian pyramid)
  float* blend_out;       // optional (debug dump of the blended pyramid)
  float* r_out;           // float output of level k, or nullptr
  uint8_t* u8_out;        // uint8 output (canvas), or nullptr
  size_t u8_pitch;
  int u8_w, u8_h;         // canvas bounds (copyRoiKernelBatched checks)
  int ox, oy;             // canvas position of level pixel (0,0)
  int hi_w, hi_h, lo_w, lo_h;
  int win_x, win_y, win_w, win_h;  // pixels of level k evaluated
  int tiles_x, tiles_y;
};

template <typename THi, int C, int N, int KIND>
__global__ void __launch_bounds__(256) up_kernel(UpArgs<THi, N> a) {
  const TileXY p0 = tile_pixel(a.tiles_x, a.tiles_y);
  const int x = a.win_x + p0.x;
  const int y = a.win_y + p0.y;
  const int b = blockIdx.z;
  if (x >= a.win_x + a.win_w || y >= a.win_y + a.win_h) return;
  const size_t hpix = ((size_t)b * a.hi_h + y) * a.hi_w + x;
  const bool coarsest = a.g_lo[0] == nullptr;
  Taps t{};
  size_t lr0 = 0, lr1 = 0;
  if (!coarsest) {
    t = make_taps(x, y, a.lo_w, a.lo_h);
    lr0 = ((size_t)b * a.lo_h + t.y0) * a.lo_w;
    lr1 = ((size_t)b * a.lo_h + t.y1) * a.lo_w;
  }
  float4 L[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const float4 g = ld_px<C>(a.g_hi[i], hpix);
    if (coarsest) {
      L[i] = g;
    } else {
      const float4 up = up_lap<C>(a.g_lo[i], lr0, lr1, t);
      L[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
    }
    if (a.lap_out[i]) st_px<C>(a.lap_out[i], hpix, L[i]);
  }
  const size_t mpix = (size_t)y * a.hi_w + x;
  float4 B;
  if constexpr (KIND == kBlend2) {
    B = blend2_px<C>(L[0], L[1], a.mask[mpix]);
  } else if constexpr (KIND == kBlend3) {
    B = blend3_px<C>(L, a.mask + mpix * 3);
  } else {
    B = blendN_px<C, N>(L, a.mask + mpix * N);
  }
  if (a.blend_out) st_px<C>(a.blend_out, hpix, B);
  float4 R = B;
  if (a.r_lo) {
    const float4 up = up_alpha<C>(a.r_lo, lr0, lr1, t);
    R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
  }
  if (a.r_out) st_px<C>(a.r_out, hpix, R);
  if (a.u8_out) {
    const int X = x + a.ox, Y = y + a.oy;
    if (X < a.u8_w && Y < a.u8_h) {
      uint8_t* p = a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C;
      st_px_u8<C>(p, R);
    }
  }
}

}  // namespace hmdev
