// hm_pyramid_fused.hpp -- whole-pyramid kernels for gfx950.
//
// down_fused_kernel: one workgroup turns a 32x32 level-0 tile into the
//   matching 16x16 .. 1x1 tiles of up to five coarser levels in LDS and writes
//   them once.  The 2x2 mean never looks outside its own block, so aligned
//   tiles need no halo and the result is the per-level kernel's exactly.
// up_fused_kernel: one workgroup reconstructs a 64x64 level-0 tile.  Going
//   from the coarsest level down, it evaluates the Laplacian of every image,
//   the blend and the reconstruction on the tile's footprint at each level
//   (plus the one-pixel halo the top-left aligned upsample needs), keeping
//   R_k in LDS; only level 0 is written, straight to the caller's output.
//   Level 0 is processed in 2x2 quads, whose four pixels share their taps.
// Per-pixel arithmetic is the same device code as the per-level kernels
// (hm_pyramid_kernels.hpp), in the same order.
#pragma once

#include "hm_pyramid_kernels.hpp"

namespace hmdev {

// Per-level buffers of levels >= 1 (device-resident table).
struct LevelDesc {
  const float* g[kMaxImages];  // Gaussian level of each image
  float* g_out[kMaxImages];    // same buffers, writable (down pass)
  const float* mask;           // mask level, M interleaved channels
  int w, h;
};

// ---------------------------------------------------------------------------
// tap-value forms of the upsample helpers
// ---------------------------------------------------------------------------
template <int C>
__device__ __forceinline__ float4 up_alpha_vals(const float4& p00, const float4& p10, const float4& p01,
                                                const float4& p11, float dx, float dy) {
  const float w00 = (1.0f - dx) * (1.0f - dy);
  const float w10 = dx * (1.0f - dy);
  const float w01 = (1.0f - dx) * dy;
  const float w11 = dx * dy;
  float sw = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  if (C != 4 || p00.w != 0.0f) { sw += w00; sx += p00.x * w00; sy += p00.y * w00; sz += p00.z * w00; }
  if (C != 4 || p10.w != 0.0f) { sw += w10; sx += p10.x * w10; sy += p10.y * w10; sz += p10.z * w10; }
  if (C != 4 || p01.w != 0.0f) { sw += w01; sx += p01.x * w01; sy += p01.y * w01; sz += p01.z * w01; }
  if (C != 4 || p11.w != 0.0f) { sw += w11; sx += p11.x * w11; sy += p11.y * w11; sz += p11.z * w11; }
  if (sw > 0.0f) return make_float4(div_exact(sx, sw), div_exact(sy, sw), div_exact(sz, sw), 0.0f);
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

template <int C>
__device__ __forceinline__ float4 up_lap_vals(const float4& p00, const float4& p10, const float4& p01,
                                              const float4& p11, float dx, float dy) {
  if constexpr (C == 4) {
    return up_alpha_vals<4>(p00, p10, p01, p11, dx, dy);
  } else {
    return make_float4(lerp2(p00.x, p10.x, p01.x, p11.x, dx, dy), lerp2(p00.y, p10.y, p01.y, p11.y, dx, dy),
                       lerp2(p00.z, p10.z, p01.z, p11.z, dx, dy), 0.0f);
  }
}

template <int C, int N, int KIND>
__device__ __forceinline__ float4 blend_any(const float4* L, const float* mask, size_t mpix) {
  if constexpr (KIND == kBlend2) {
    return blend2_px<C>(L[0], L[1], mask[mpix]);
  } else if constexpr (KIND == kBlend3) {
    return blend3_px<C>(L, mask + mpix * 3);
  } else {
    return blendN_px<C, N>(L, mask + mpix * N);
  }
}

// ---------------------------------------------------------------------------
// down pass
// ---------------------------------------------------------------------------
constexpr int kDownTile = 32;  // level-0 tile edge; up to 5 coarser levels per launch

template <typename T0, int N>
struct DownFusedArgs {
  const T0* g0[N];          // level 0 (tight, batch-major)
  int w0, h0;
  int n_out;                // levels produced (1..5)
  const LevelDesc* lv;      // lv[0] describes the first produced level, lv[1] the next ...
  int tiles_x, tiles_y;
};

template <typename T0, int C, int N>
__global__ void __launch_bounds__(256) down_fused_kernel(DownFusedArgs<T0, N> a) {
  __shared__ float4 s1[N][16 * 16];
  __shared__ float4 s2[N][8 * 8];
  const int t = xcd_tile((int)blockIdx.x, a.tiles_x * a.tiles_y);
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blockIdx.z;
  const int tid = threadIdx.x;
  // ---- first produced level: one pixel per thread from level 0
  {
    const LevelDesc& d = a.lv[0];
    const int lx = tid & 15, ly = tid >> 4;
    const int x = tx * 16 + lx, y = ty * 16 + ly;
    if (x < d.w && y < d.h) {
      const int ix = 2 * x, iy = 2 * y;
      const bool has_x1 = ix + 1 < a.w0, has_y1 = iy + 1 < a.h0;
      const size_t r0 = ((size_t)b * a.h0 + iy) * a.w0;
      const size_t r1 = r0 + a.w0;
      float4 p[N][4];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        p[i][0] = ld_px<C>(a.g0[i], r0 + ix);
        if (has_x1) p[i][1] = ld_px<C>(a.g0[i], r0 + ix + 1);
        if (has_y1) p[i][2] = ld_px<C>(a.g0[i], r1 + ix);
        if (has_x1 && has_y1) p[i][3] = ld_px<C>(a.g0[i], r1 + ix + 1);
      }
      const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
      const size_t o = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float4 r = down_px<C>(p[i], in_range);
        s1[i][ly * 16 + lx] = r;
        st_px<C>(d.g_out[i], o, r);
      }
    }
  }
  // ---- further levels from LDS: 8x8, 4x4, 2x2, 1x1 per tile
  int edge = 16;
  float4(*src)[16 * 16] = s1;
  for (int k = 1; k < a.n_out; ++k) {
    __syncthreads();
    const LevelDesc& din = a.lv[k - 1];
    const LevelDesc& d = a.lv[k];
    const int oedge = edge >> 1;
    const int x0 = tx * oedge, y0 = ty * oedge;
    float4 r[N];
    bool active = false;
    int lx = 0, ly = 0;
    if (tid < oedge * oedge) {
      lx = tid % oedge;
      ly = tid / oedge;
      const int x = x0 + lx, y = y0 + ly;
      if (x < d.w && y < d.h) {
        active = true;
        const int gx = 2 * x, gy = 2 * y;  // level k-1 coordinates
        const bool has_x1 = gx + 1 < din.w, has_y1 = gy + 1 < din.h;
        const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
        const int sx = 2 * lx, sy = 2 * ly;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          float4 p[4];
          p[0] = (k == 1 ? src[i] : s2[i])[sy * edge + sx];
          if (has_x1) p[1] = (k == 1 ? src[i] : s2[i])[sy * edge + sx + 1];
          if (has_y1) p[2] = (k == 1 ? src[i] : s2[i])[(sy + 1) * edge + sx];
          if (has_x1 && has_y1) p[3] = (k == 1 ? src[i] : s2[i])[(sy + 1) * edge + sx + 1];
          r[i] = down_px<C>(p, in_range);
          st_px<C>(d.g_out[i], ((size_t)b * d.h + y) * d.w + x, r[i]);
        }
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int i = 0; i < N; ++i) s2[i][ly * oedge + lx] = r[i];
    }
    edge = oedge;
  }
}

// ---------------------------------------------------------------------------
// up pass
// ---------------------------------------------------------------------------
constexpr int kUpTile = 64;                        // level-0 tile edge
constexpr int kUpRegion = kUpTile / 2 + 2;         // max footprint edge at levels >= 1

template <typename T0, int N>
struct UpFusedArgs {
  const T0* g0[N];            // level-0 images
  const float* mask0;         // level-0 mask
  int w0, h0;
  int levels;                 // total levels (>= 1)
  const LevelDesc* lv;        // lv[k - 1] describes level k
  float* f32_out;             // level-0 float output (tight), or nullptr
  uint8_t* u8_out;            // level-0 uint8 canvas, or nullptr
  size_t u8_pitch;
  int u8_w, u8_h, ox, oy;
  int win_x, win_y, win_w, win_h;  // level-0 pixels produced
  int tiles_x, tiles_y;
};

template <typename T0, int C, int N, int KIND>
__global__ void __launch_bounds__(256) up_fused_kernel(UpFusedArgs<T0, N> a) {
  __shared__ float4 buf[2][kUpRegion * kUpRegion];
  __shared__ int reg[kMaxLevels][4];  // footprint [x0, x1] x [y0, y1] per level
  const int t = xcd_tile((int)blockIdx.x, a.tiles_x * a.tiles_y);
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blockIdx.z;
  const int tid = threadIdx.x;
  const int L = a.levels;
  const int ax0 = a.win_x + tx * kUpTile, ay0 = a.win_y + ty * kUpTile;
  const int bx0 = min(ax0 + kUpTile, a.win_x + a.win_w) - 1;
  const int by0 = min(ay0 + kUpTile, a.win_y + a.win_h) - 1;
  if (tid == 0) {
    int x0 = ax0, x1 = bx0, y0 = ay0, y1 = by0;
    reg[0][0] = x0; reg[0][1] = x1; reg[0][2] = y0; reg[0][3] = y1;
    for (int k = 1; k < L; ++k) {
      const LevelDesc& d = a.lv[k - 1];
      x0 >>= 1;
      y0 >>= 1;
      x1 = min((x1 >> 1) + 1, d.w - 1);
      y1 = min((y1 >> 1) + 1, d.h - 1);
      reg[k][0] = x0; reg[k][1] = x1; reg[k][2] = y0; reg[k][3] = y1;
    }
  }
  __syncthreads();

  // ---- levels L-1 .. 1: R_k on the footprint, in LDS
  int cur = 0;
  for (int k = L - 1; k >= 1; --k) {
    const LevelDesc& d = a.lv[k - 1];
    const int rx0 = reg[k][0], ry0 = reg[k][2];
    const int rw = reg[k][1] - rx0 + 1, rh = reg[k][3] - ry0 + 1;
    const bool coarsest = k == L - 1;
    int nx0 = 0, ny0 = 0, nrw = 1;
    const LevelDesc* dn = nullptr;
    if (!coarsest) {
      dn = &a.lv[k];
      nx0 = reg[k + 1][0];
      ny0 = reg[k + 1][2];
      nrw = reg[k + 1][1] - nx0 + 1;
    }
    for (int idx = tid; idx < rw * rh; idx += blockDim.x) {
      const int x = rx0 + idx % rw, y = ry0 + idx / rw;
      const size_t pix = ((size_t)b * d.h + y) * d.w + x;
      float4 Lv[N];
      float4 R;
      if (coarsest) {
#pragma unroll
        for (int i = 0; i < N; ++i) Lv[i] = ld_px<C>(d.g[i], pix);
        R = blend_any<C, N, KIND>(Lv, d.mask, (size_t)y * d.w + x);
      } else {
        const int qx = x >> 1, qy = y >> 1;
        const int qx1 = min(qx + 1, dn->w - 1), qy1 = min(qy + 1, dn->h - 1);
        const float fdx = (x & 1) ? 0.5f : 0.0f, fdy = (y & 1) ? 0.5f : 0.0f;
        const size_t l0 = ((size_t)b * dn->h + qy) * dn->w, l1 = ((size_t)b * dn->h + qy1) * dn->w;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float4 g = ld_px<C>(d.g[i], pix);
          const float4 up = up_lap_vals<C>(ld_px<C>(dn->g[i], l0 + qx), ld_px<C>(dn->g[i], l0 + qx1),
                                           ld_px<C>(dn->g[i], l1 + qx), ld_px<C>(dn->g[i], l1 + qx1), fdx, fdy);
          Lv[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        }
        const float4 B = blend_any<C, N, KIND>(Lv, d.mask, (size_t)y * d.w + x);
        const float4* prev = buf[cur ^ 1];
        const int lx = qx - nx0, lx1 = qx1 - nx0, ly = qy - ny0, ly1 = qy1 - ny0;
        const float4 up = up_alpha_vals<C>(prev[ly * nrw + lx], prev[ly * nrw + lx1], prev[ly1 * nrw + lx],
                                           prev[ly1 * nrw + lx1], fdx, fdy);
        R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
      }
      buf[cur][idx] = R;
    }
    __syncthreads();
    cur ^= 1;
  }

  // ---- level 0 in 2x2 quads sharing their level-1 taps
  const int qx0 = ax0 >> 1, qy0 = ay0 >> 1;
  const int nqx = (bx0 >> 1) - qx0 + 1, nqy = (by0 >> 1) - qy0 + 1;
  const float4* prev = buf[cur ^ 1];
  int nx0 = 0, ny0 = 0, nrw = 1;
  const LevelDesc* dn = nullptr;
  if (L > 1) {
    dn = &a.lv[0];
    nx0 = reg[1][0];
    ny0 = reg[1][2];
    nrw = reg[1][1] - nx0 + 1;
  }
  for (int q = tid; q < nqx * nqy; q += blockDim.x) {
    const int qx = qx0 + q % nqx, qy = qy0 + q / nqx;
    float4 gt[N][4];
    float4 rt[4];
    if (L > 1) {
      const int qx1 = min(qx + 1, dn->w - 1), qy1 = min(qy + 1, dn->h - 1);
      const size_t l0 = ((size_t)b * dn->h + qy) * dn->w, l1 = ((size_t)b * dn->h + qy1) * dn->w;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        gt[i][0] = ld_px<C>(dn->g[i], l0 + qx);
        gt[i][1] = ld_px<C>(dn->g[i], l0 + qx1);
        gt[i][2] = ld_px<C>(dn->g[i], l1 + qx);
        gt[i][3] = ld_px<C>(dn->g[i], l1 + qx1);
      }
      const int lx = qx - nx0, lx1 = qx1 - nx0, ly = qy - ny0, ly1 = qy1 - ny0;
      rt[0] = prev[ly * nrw + lx];
      rt[1] = prev[ly * nrw + lx1];
      rt[2] = prev[ly1 * nrw + lx];
      rt[3] = prev[ly1 * nrw + lx1];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
      if (x < ax0 || x > bx0 || y < ay0 || y > by0) continue;
      const float fdx = (e & 1) ? 0.5f : 0.0f, fdy = (e >> 1) ? 0.5f : 0.0f;
      const size_t pix = ((size_t)b * a.h0 + y) * a.w0 + x;
      float4 Lv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float4 g = ld_px<C>(a.g0[i], pix);
        if (L > 1) {
          const float4 up = up_lap_vals<C>(gt[i][0], gt[i][1], gt[i][2], gt[i][3], fdx, fdy);
          Lv[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        } else {
          Lv[i] = g;
        }
      }
      const float4 B = blend_any<C, N, KIND>(Lv, a.mask0, (size_t)y * a.w0 + x);
      float4 R = B;
      if (L > 1) {
        const float4 up = up_alpha_vals<C>(rt[0], rt[1], rt[2], rt[3], fdx, fdy);
        R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
      }
      if (a.f32_out) st_px<C>(a.f32_out, pix, R);
      if (a.u8_out) {
        const int X = x + a.ox, Y = y + a.oy;
        if (X < a.u8_w && Y < a.u8_h)
          st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
      }
    }
  }
}

}  // namespace hmdev
