// hm_pyramid_fused.hpp -- whole-pyramid kernels for gfx950.
//
// down_fused_kernel: one workgroup turns a 32x32 level-0 tile into the
//   matching 16x16 .. 1x1 tiles of up to five coarser levels in LDS and writes
//   them once.  The 2x2 mean never looks outside its own block, so aligned
//   tiles need no halo and the result is the per-level kernel's exactly.
// up_quad_kernel: one level, one thread per 2x2 quad: the Laplacian of every
//   image from the level's pixels and the shared taps of the next level, the
//   blend, and the reconstruction from the shared R taps of the next level.
//   Every level of the reconstruction runs it, coarsest first; it keeps no
//   state, so it runs at full occupancy, and level 0 writes straight to the
//   output.  The four pixels of a quad share their taps and sit at fixed
//   upsample positions (dx, dy) in {0, 1/2}^2.
// Per-pixel arithmetic is the same as the per-level kernels
// (hm_pyramid_kernels.hpp), in the same order.
#pragma once

#include <type_traits>

#include "hm_pyramid_kernels.hpp"

namespace hmdev {

// Inclusive pixel rectangle; empty when x1 < x0 or y1 < y0.
struct ImgRect {
  int x0, y0, x1, y1;
};
__host__ __device__ __forceinline__ bool in_rect(const ImgRect& r, int x, int y) {
  return x >= r.x0 && x <= r.x1 && y >= r.y0 && y <= r.y1;
}

// Per-level buffers of levels >= 1.
//
// Sparse images: outside its support s[i] an image's Gaussian level is
// exactly zero (alpha 0), because its level-0 input is zero outside the
// image's box and a 2x2 mean of zeros is zero.  The buffers hold valid values
// only on z[i] (the support grown by 2 pixels): there the down pass writes
// them (zero outside s[i]), and only there does the up pass read them; at a
// pixel outside z[i] the image's Laplacian is exactly zero (its level value
// and every upsample tap are zero), which the up pass uses without loading.
// Dense pyramids have s[i] = z[i] = the whole level.
struct LevelDesc {
  const float* g[kMaxImages];  // Gaussian level of each image
  float* g_out[kMaxImages];    // same buffers, writable (down pass)
  const float* mask;           // mask level, M interleaved channels
  int w, h;
  ImgRect s[kMaxImages];       // support of each image
  ImgRect z[kMaxImages];       // valid region of each image's buffer
};

// ---------------------------------------------------------------------------
// Upsample of the pixels of a 2x2 quad from its four shared taps.
// Pixel e of the quad is (e & 1, e >> 1) and sits at the fixed fractional
// position (dx, dy) = ((e & 1) / 2, (e >> 1) / 2) between taps 00, 10, 01,
// 11.  These forms give the same bits as the general bilinear with those
// constants (cudaBlend.cu:416-449, :649-666): a tap of weight 0 adds +0 or -0
// to a running sum that starts at +0 and, rounding to nearest, never becomes
// -0, so skipping it never changes the sum; a tap of weight 1 is its own
// product; and dividing by a power-of-two weight sum is exact.  Values can
// differ from the general form only in the sign of a zero, which no later
// step (comparisons, sums with non-zero terms, conversion to uint8) can see.
// ---------------------------------------------------------------------------
// Orders one wavefront's LDS accesses where its lanes overwrite entries that
// other lanes of the same wavefront have just read (or read what they have
// just written).  The hardware executes a wavefront's LDS instructions in
// order; the barrier keeps the scheduler from moving instructions across it,
// and the wavefront-scope fence (no instruction) keeps the compiler from
// reordering the memory accesses themselves.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float4 rgb_of(const float4& p) { return make_float4(p.x, p.y, p.z, 0.0f); }
__device__ __forceinline__ float4 half_pair(const float4& p, const float4& q) {
  return make_float4(p.x * 0.5f + q.x * 0.5f, p.y * 0.5f + q.y * 0.5f, p.z * 0.5f + q.z * 0.5f, 0.0f);
}
// Two taps of weight 1/2 each; a lone valid tap has weight sum 1/2 and gives
// (v * 0.5) * 2 == v.
__device__ __forceinline__ float4 pair_alpha(bool a, const float4& p, bool b, const float4& q) {
  if (a && b) return half_pair(p, q);
  if (a) return rgb_of(p);
  if (b) return rgb_of(q);
  return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Alpha-aware form; ALPHA = false takes every tap (3-channel reconstruct).
template <bool ALPHA>
__device__ __forceinline__ float4 up_quad_alpha(int e, const float4& p00, const float4& p10, const float4& p01,
                                                const float4& p11) {
  const bool a00 = !ALPHA || p00.w != 0.0f, a10 = !ALPHA || p10.w != 0.0f;
  const bool a01 = !ALPHA || p01.w != 0.0f, a11 = !ALPHA || p11.w != 0.0f;
  if (e == 0) return a00 ? rgb_of(p00) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (e == 1) return pair_alpha(a00, p00, a10, p10);
  if (e == 2) return pair_alpha(a00, p00, a01, p01);
  // every tap with weight 1/4, summed in the order 00, 10, 01, 11
  float sx = 0.0f, sy = 0.0f, sz = 0.0f;
  int n = 0;
  if (a00) { sx += p00.x * 0.25f; sy += p00.y * 0.25f; sz += p00.z * 0.25f; ++n; }
  if (a10) { sx += p10.x * 0.25f; sy += p10.y * 0.25f; sz += p10.z * 0.25f; ++n; }
  if (a01) { sx += p01.x * 0.25f; sy += p01.y * 0.25f; sz += p01.z * 0.25f; ++n; }
  if (a11) { sx += p11.x * 0.25f; sy += p11.y * 0.25f; sz += p11.z * 0.25f; ++n; }
  if (n == 3) return make_float4(sx / 0.75f, sy / 0.75f, sz / 0.75f, 0.0f);
  const float k = n == 1 ? 4.0f : (n == 2 ? 2.0f : (n == 4 ? 1.0f : 0.0f));  // n == 0: +0 * 0
  return make_float4(sx * k, sy * k, sz * k, 0.0f);
}

// The plain two-stage lerp of the 3-channel Laplacian (cudaBlend.cu:446-449).
__device__ __forceinline__ float4 up_quad_lerp(int e, const float4& p00, const float4& p10, const float4& p01,
                                               const float4& p11) {
  if (e == 0) return rgb_of(p00);
  if (e == 1) return half_pair(p00, p10);
  if (e == 2) return half_pair(p00, p01);
  return half_pair(half_pair(p00, p10), half_pair(p01, p11));
}

// Upsample of the Laplacian (alpha-aware for RGBA, plain for RGB) and of the
// reconstruction (alpha-aware for RGBA, every tap for RGB).
template <int C>
__device__ __forceinline__ float4 up_quad_lap(int e, const float4* p) {
  if constexpr (C == 4)
    return up_quad_alpha<true>(e, p[0], p[1], p[2], p[3]);
  else
    return up_quad_lerp(e, p[0], p[1], p[2], p[3]);
}
template <int C>
__device__ __forceinline__ float4 up_quad_rec(int e, const float4* p) {
  return up_quad_alpha<C == 4>(e, p[0], p[1], p[2], p[3]);
}

// How the level-0 mask is stored.  The values are the same floats the
// reference's float mask holds: a uint8 seam converted to float is exact, and
// a one-hot mask is 1.0 for the pixel's label and 0.0 elsewhere.
enum Mask0Mode { kMask0F32 = 0, kMask0U8 = 1, kMask0Label = 2 };

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

// One produced level, carried by value in the kernel arguments (constant
// memory: uniform scalar loads, instead of vector loads of a global table the
// kernel also writes through).
template <int N>
struct DownLevel {
  float* g_out[N];
  ImgRect s[N], z[N];
  int w, h;
};

template <typename T0, int N>
struct DownFusedArgs {
  const T0* g0[N];          // input level (tight, batch-major)
  int w0, h0;
  int n_out;                // levels produced (1..5)
  DownLevel<N> lv[5];       // lv[0] describes the first produced level, lv[1] the next ...
  int tiles_x, tiles_y;
  const uint32_t* touch;    // three or more images: per tile, the images whose valid
                            // region meets it at some produced level (bit i), or nullptr
};

// Levels 2 .. n_out of the down pass: the first produced level's 16x16 tile
// of every image is in s1; each further level's tile (8x8, 4x4, 2x2, 1x1)
// comes from the one above it and replaces it at the start of s1 (each
// level's reads are done before its writes), and is written once.  These levels fit
// one wavefront (64 pixels at most), so after the barrier that completes s1
// the other three wavefronts end, freeing their slots for the next tiles'
// loads, and wavefront 0 runs the chain alone: a wavefront's LDS operations
// complete in program order, so within it each level's reads of the level
// above are done before its writes replace them, and its writes before the
// next level's reads.  Must be the caller's last step (no barrier follows).
template <int C, int N, typename Args>
__device__ __forceinline__ void down_rest(const Args& a, int tx, int ty, int b, int tid, float4 (*s1)[16 * 16]) {
  if (a.n_out < 2) return;
  __syncthreads();  // the first produced level's tile is complete
  if (tid >= 64) return;
  int edge = 16;
  for (int k = 1; k < a.n_out; ++k) {
    const DownLevel<N>& din = a.lv[k - 1];
    const DownLevel<N>& d = a.lv[k];
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
          const float4* src = s1[i];
          float4 p[4];
          p[0] = src[sy * edge + sx];
          if (has_x1) p[1] = src[sy * edge + sx + 1];
          if (has_y1) p[2] = src[(sy + 1) * edge + sx];
          if (has_x1 && has_y1) p[3] = src[(sy + 1) * edge + sx + 1];
          r[i] = down_px<C>(p, in_range);  // zero outside the support: its LDS inputs are zero
          if (in_rect(d.z[i], x, y)) st_px<C>(d.g_out[i], ((size_t)b * d.h + y) * d.w + x, r[i]);
        }
      }
    }
    wave_lds_order();  // the level's reads stay before its writes
    if (active) {
#pragma unroll
      for (int i = 0; i < N; ++i) s1[i][ly * oedge + lx] = r[i];
    }
    wave_lds_order();  // and its writes before the next level's reads
    edge = oedge;
  }
}

// The host's fill of DownFusedArgs from the pyramid's level table.
template <typename T0, int N>
inline void fill_down_args(DownFusedArgs<T0, N>& a, const LevelDesc* lv, int w0, int h0, int n_out) {
  a.w0 = w0;
  a.h0 = h0;
  a.n_out = n_out;
  for (int k = 0; k < 5; ++k) {
    const LevelDesc& d = lv[k < n_out ? k : 0];
    for (int i = 0; i < N; ++i) {
      a.lv[k].g_out[i] = d.g_out[i];
      a.lv[k].s[i] = d.s[i];
      a.lv[k].z[i] = d.z[i];
    }
    a.lv[k].w = d.w;
    a.lv[k].h = d.h;
  }
  a.tiles_x = (w0 + kDownTile - 1) / kDownTile;
  a.tiles_y = (h0 + kDownTile - 1) / kDownTile;
  a.touch = nullptr;
}

// LDS of the down pass: the first produced level's 16x16 tile of every image
// (or slot); the further levels' tiles take its place.
template <int N>
constexpr int down_lds_float4() { return N * 16 * 16; }

template <typename T0, int C, int N>
__device__ __forceinline__ void down_fused_body(const DownFusedArgs<T0, N>& a, Blk blk, float4* lds) {
  float4(*s1)[16 * 16] = reinterpret_cast<float4(*)[16 * 16]>(lds);
  const int t = xcd_tile(blk.x, a.tiles_x * a.tiles_y);
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blk.z;
  const int tid = flat_tid();
  // ---- first produced level: one pixel per thread from level 0
  {
    const DownLevel<N>& d = a.lv[0];
    const int lx = tid & 15, ly = tid >> 4;
    const int x = tx * 16 + lx, y = ty * 16 + ly;
    if (x < d.w && y < d.h) {
      const int ix = 2 * x, iy = 2 * y;
      const bool has_x1 = ix + 1 < a.w0, has_y1 = iy + 1 < a.h0;
      const size_t r0 = ((size_t)b * a.h0 + iy) * a.w0;
      const size_t r1 = r0 + a.w0;
      float4 p[N][4];
      bool in_s[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        // Outside the image's support the result is zero (one or two images
        // are dense, so the support is the whole level).  Taps past the edge
        // load their in-range neighbour, ignored by down_px: no branch between
        // the loads, so all of them are in flight together.
        in_s[i] = in_rect(d.s[i], x, y);
        const int ix1 = has_x1 ? ix + 1 : ix;
        const size_t rr1 = has_y1 ? r1 : r0;
        p[i][0] = ld_px<C>(a.g0[i], r0 + ix);
        p[i][1] = ld_px<C>(a.g0[i], r0 + ix1);
        p[i][2] = ld_px<C>(a.g0[i], rr1 + ix);
        p[i][3] = ld_px<C>(a.g0[i], rr1 + ix1);
      }
      const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
      const size_t o = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const float4 r = in_s[i] ? down_px<C>(p[i], in_range) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        s1[i][ly * 16 + lx] = r;
        if (in_rect(d.z[i], x, y)) st_px<C>(d.g_out[i], o, r);
      }
    }
  }
  down_rest<C, N>(a, tx, ty, b, tid, s1);
}

// Three or more images: the images whose valid region meets the tile at some
// produced level run the level chain together, up to kDownSlots at a time in
// one set of LDS tiles; a workgroup skips the other images (their buffers
// there are never read).  The slot count S is a template parameter (the
// kernel switches on it, a uniform value), so all loads of the first level
// are issued without a branch between them.  Two slots: on the 8-image
// workload (14000x2600 canvas) the pass takes 107 us, against 121 us with
// one slot, 116 us with three and 130 us with four; three images on a
// 6200x2300 canvas: 28.6 us against 30.8 us with three slots.
constexpr int kDownSlots = 2;

// Whether the e x e tile (tx, ty) of a level meets the rectangle z.
__host__ __device__ __forceinline__ bool tile_meets(const ImgRect& z, int tx, int ty, int e) {
  const int x0 = tx * e, y0 = ty * e;
  return x0 <= z.x1 && x0 + e - 1 >= z.x0 && y0 <= z.y1 && y0 + e - 1 >= z.y0;
}

// last: no chunk of the tile follows, so the levels past the first run in
// wavefront 0 alone as in down_rest (the other wavefronts end); otherwise
// the whole workgroup runs them.
template <typename T0, int C, int N, int S>
__device__ __forceinline__ void down_slots(const DownFusedArgs<T0, N>& a, const int* idx, int tx, int ty, int b,
                                           int tid, float4 (*s1)[16 * 16], bool last) {
  {
    const DownLevel<N>& d = a.lv[0];
    const int lx = tid & 15, ly = tid >> 4;
    const int x = tx * 16 + lx, y = ty * 16 + ly;
    if (x < d.w && y < d.h) {
      // Taps past the edge load their in-range neighbour (ignored by
      // down_px), and pixels outside an image's support load too (the
      // result is zero there).
      const int ix = 2 * x, iy = 2 * y;
      const bool has_x1 = ix + 1 < a.w0, has_y1 = iy + 1 < a.h0;
      const size_t r0 = ((size_t)b * a.h0 + iy) * a.w0;
      const size_t rr1 = has_y1 ? r0 + a.w0 : r0;
      const int ix1 = has_x1 ? ix + 1 : ix;
      float4 p[S][4];
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const T0* g = a.g0[idx[s]];
        p[s][0] = ld_px<C>(g, r0 + ix);
        p[s][1] = ld_px<C>(g, r0 + ix1);
        p[s][2] = ld_px<C>(g, rr1 + ix);
        p[s][3] = ld_px<C>(g, rr1 + ix1);
      }
      const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
      const size_t o = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int i = idx[s];
        const float4 r = in_rect(d.s[i], x, y) ? down_px<C>(p[s], in_range) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        s1[s][ly * 16 + lx] = r;
        if (in_rect(d.z[i], x, y)) st_px<C>(d.g_out[i], o, r);
      }
    }
  }
  if (a.n_out < 2) return;
  __syncthreads();  // the first produced level's tiles are complete
  if (last && tid >= 64) return;
  int edge = 16;
  for (int k = 1; k < a.n_out; ++k) {
    const DownLevel<N>& din = a.lv[k - 1];
    const DownLevel<N>& d = a.lv[k];
    const int oedge = edge >> 1;
    float4 r[S];
    bool active = false;
    int lx = 0, ly = 0;
    if (tid < oedge * oedge) {
      lx = tid % oedge;
      ly = tid / oedge;
      const int x = tx * oedge + lx, y = ty * oedge + ly;
      if (x < d.w && y < d.h) {
        active = true;
        const int gx = 2 * x, gy = 2 * y;
        const bool has_x1 = gx + 1 < din.w, has_y1 = gy + 1 < din.h;
        const bool in_range[4] = {true, has_x1, has_y1, has_x1 && has_y1};
        const int sx = 2 * lx, sy = 2 * ly;
        const size_t o = ((size_t)b * d.h + y) * d.w + x;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float4* src = s1[s];
          float4 p[4];
          p[0] = src[sy * edge + sx];
          if (has_x1) p[1] = src[sy * edge + sx + 1];
          if (has_y1) p[2] = src[(sy + 1) * edge + sx];
          if (has_x1 && has_y1) p[3] = src[(sy + 1) * edge + sx + 1];
          r[s] = down_px<C>(p, in_range);  // zero outside the support: its LDS inputs are zero
          const int i = idx[s];
          if (in_rect(d.z[i], x, y)) st_px<C>(d.g_out[i], o, r[s]);
        }
      }
    }
    if (last)
      wave_lds_order();  // the level's reads stay before its writes
    else
      __syncthreads();
    if (active) {
#pragma unroll
      for (int s = 0; s < S; ++s) s1[s][ly * oedge + lx] = r[s];
    }
    if (last)
      wave_lds_order();  // and its writes before the next level's reads
    else
      __syncthreads();
    edge = oedge;
  }
}

template <int N>
constexpr int down_slots_count() { return N < kDownSlots ? N : kDownSlots; }

template <typename T0, int C, int N>
__device__ __forceinline__ void down_fused_slots_body(const DownFusedArgs<T0, N>& a, Blk blk, float4* lds) {
  constexpr int K = down_slots_count<N>();
  float4(*s1)[16 * 16] = reinterpret_cast<float4(*)[16 * 16]>(lds);
  const int t = xcd_tile(blk.x, a.tiles_x * a.tiles_y);
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blk.z;
  const int tid = flat_tid();
  unsigned act = 0;  // uniform
  if (a.touch) {
    act = __builtin_amdgcn_readfirstlane(a.touch[t]);
  } else {
#pragma unroll
    for (int i = 0; i < N; ++i) {
      for (int k = 0; k < a.n_out; ++k) {
        const ImgRect& z = a.lv[k].z[i];
        if (tile_meets(z, tx, ty, 16 >> k)) act |= 1u << i;
      }
    }
  }
  bool first = true;
  while (act) {
    int idx[K];
    const int n = min(__builtin_popcount(act), K);
#pragma unroll
    for (int s = 0; s < K; ++s) {
      idx[s] = act ? __builtin_ctz(act) : 0;
      act &= act - 1u;
    }
    if (!first) __syncthreads();  // the previous slots' LDS reads are done
    first = false;
    const bool last = act == 0;  // uniform
    switch (n) {  // uniform
      case 1: down_slots<T0, C, N, 1>(a, idx, tx, ty, b, tid, s1, last); break;
      case 2:
        if constexpr (K >= 2) down_slots<T0, C, N, 2>(a, idx, tx, ty, b, tid, s1, last);
        break;
      case 3:
        if constexpr (K >= 3) down_slots<T0, C, N, 3>(a, idx, tx, ty, b, tid, s1, last);
        break;
      default:
        if constexpr (K >= 4) down_slots<T0, C, N, 4>(a, idx, tx, ty, b, tid, s1, last);
        break;
    }
  }
}

// Every down launch's body: all images (one or two), or the images of the
// tile in slots.
template <typename T0, int C, int N>
__device__ __forceinline__ void down_any_body(const DownFusedArgs<T0, N>& a, Blk blk, float4* lds) {
  if constexpr (N <= 2)
    down_fused_body<T0, C, N>(a, blk, lds);
  else
    down_fused_slots_body<T0, C, N>(a, blk, lds);
}

template <typename T0, int C, int N>
__global__ void __launch_bounds__(256) down_fused_kernel(DownFusedArgs<T0, N> a) {
  __shared__ float4 lds[down_lds_float4<down_slots_count<N>()>()];
  (void)a;
  down_any_body<T0, C, N>(kernel_arg<DownFusedArgs<T0, N>>(), hw_blk(), lds);
}

// ---------------------------------------------------------------------------
// up pass
// ---------------------------------------------------------------------------
template <int C, int N, int KIND, int MODE>
__device__ __forceinline__ float4 blend_level0(const float4* L, const void* mask, size_t mpix) {
  if constexpr (MODE == kMask0F32) {
    return blend_any<C, N, KIND>(L, static_cast<const float*>(mask), mpix);
  } else if constexpr (MODE == kMask0U8) {
    static_assert(KIND == kBlend2, "a one-channel uint8 mask belongs to the 2-image blend");
    return blend2_px<C>(L[0], L[1], (float)static_cast<const uint8_t*>(mask)[mpix]);
  } else {
    static_assert(KIND != kBlend2, "label masks belong to the 3- and N-image blends");
    const int label = static_cast<const uint8_t*>(mask)[mpix];
    float mp[N];
#pragma unroll
    for (int i = 0; i < N; ++i) mp[i] = (label == i) ? 1.0f : 0.0f;
    if constexpr (KIND == kBlend3)
      return blend3_px<C>(L, mp);
    else
      return blendN_px<C, N>(L, mp);
  }
}

// ---------------------------------------------------------------------------
// the small tail of the up pass, one launch (one or two images)
// ---------------------------------------------------------------------------
// Levels base .. L-1 are small (1/16 of level 0 and less).  One per-level
// launch each would be a chain of latency-bound launches; instead one
// workgroup reconstructs a kTailTile x kTailTile tile of the base level,
// going from the coarsest level down over the tile's footprint at each level
// (plus the one-pixel halo of the top-left aligned upsample), with R of the
// level below the current one in LDS.  Only R_base is written.  Every level's
// pixels are evaluated in 2x2 quads sharing their taps.
constexpr int kTailTile = 32;
constexpr int kTailRegion = kTailTile / 2 + 2;  // footprint edge at the levels above the base
constexpr int kMaxTail = 15;

template <int N>
struct TailLevel {
  const float* g[N];  // Gaussian level of each image
  const float* mask;  // mask level, M interleaved channels
  int w, h;
  ImgRect z[N];       // valid region of each image (read by the 3- and N-image blends)
};

template <int N>
struct UpTailArgs {
  int n;                        // levels in the tail (>= 2): lv[0] = base ... lv[n-1] = coarsest
  TailLevel<N> lv[kMaxTail];
  float* r_out;                 // R of the base level (tight)
  int wx0, wy0, wx1, wy1;       // base-level pixels produced
  int tiles_x, tiles_y;
};

// LDS of the tail: R of two levels, then the footprints (as ints).
constexpr int kTailLdsFloat4 = 2 * kTailRegion * kTailRegion + kMaxTail;

template <int C, int N, int KIND>
__device__ __forceinline__ void up_tail_body(const UpTailArgs<N>& a, Blk blk, float4* lds) {
  float4(*buf)[kTailRegion * kTailRegion] = reinterpret_cast<float4(*)[kTailRegion * kTailRegion]>(lds);
  int(*reg)[4] = reinterpret_cast<int(*)[4]>(lds + 2 * kTailRegion * kTailRegion);  // footprint per tail level
  const int t = blk.x;
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blk.z;
  const int tid = flat_tid();
  const int n = a.n;
  if (tid == 0) {
    int x0 = a.wx0 + tx * kTailTile, y0 = a.wy0 + ty * kTailTile;
    int x1 = min(x0 + kTailTile - 1, a.wx1), y1 = min(y0 + kTailTile - 1, a.wy1);
    reg[0][0] = x0; reg[0][1] = x1; reg[0][2] = y0; reg[0][3] = y1;
    for (int j = 1; j < n; ++j) {
      x0 >>= 1;
      y0 >>= 1;
      x1 = min((x1 >> 1) + 1, a.lv[j].w - 1);
      y1 = min((y1 >> 1) + 1, a.lv[j].h - 1);
      reg[j][0] = x0; reg[j][1] = x1; reg[j][2] = y0; reg[j][3] = y1;
    }
  }
  __syncthreads();
  int cur = 0;
  // coarsest level: R = blend of the Gaussians
  {
    const TailLevel<N>& d = a.lv[n - 1];
    const int rx0 = reg[n - 1][0], ry0 = reg[n - 1][2];
    const int rw = reg[n - 1][1] - rx0 + 1, rh = reg[n - 1][3] - ry0 + 1;
    for (int idx = tid; idx < rw * rh; idx += kWgThreads) {
      const int x = rx0 + idx % rw, y = ry0 + idx / rw;
      const size_t pix = ((size_t)b * d.h + y) * d.w + x;
      float4 Lv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) Lv[i] = ld_px<C>(d.g[i], pix);
      buf[cur][idx] = blend_any<C, N, KIND>(Lv, d.mask, (size_t)y * d.w + x);
    }
  }
  // levels n-2 .. 0 in quads: taps of G from global memory, of R from LDS
  for (int j = n - 2; j >= 0; --j) {
    __syncthreads();
    cur ^= 1;
    const TailLevel<N>& d = a.lv[j];
    const TailLevel<N>& dn = a.lv[j + 1];
    const int rx0 = reg[j][0], rx1 = reg[j][1], ry0 = reg[j][2], ry1 = reg[j][3];
    const int rw = rx1 - rx0 + 1;
    const int qx0 = rx0 >> 1, qy0 = ry0 >> 1;
    const int nqx = (rx1 >> 1) - qx0 + 1, nqy = (ry1 >> 1) - qy0 + 1;
    const int nx0 = reg[j + 1][0], ny0 = reg[j + 1][2], nrw = reg[j + 1][1] - nx0 + 1;
    const float4* prev = buf[cur ^ 1];
    for (int q = tid; q < nqx * nqy; q += kWgThreads) {
      const int qx = qx0 + q % nqx, qy = qy0 + q / nqx;
      const int qx1 = min(qx + 1, dn.w - 1), qy1 = min(qy + 1, dn.h - 1);
      const size_t l0 = ((size_t)b * dn.h + qy) * dn.w, l1 = ((size_t)b * dn.h + qy1) * dn.w;
      float4 gt[N][4];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        gt[i][0] = ld_px<C>(dn.g[i], l0 + qx);
        gt[i][1] = ld_px<C>(dn.g[i], l0 + qx1);
        gt[i][2] = ld_px<C>(dn.g[i], l1 + qx);
        gt[i][3] = ld_px<C>(dn.g[i], l1 + qx1);
      }
      float4 rt[4];
      const int lx = qx - nx0, lx1 = qx1 - nx0, ly = qy - ny0, ly1 = qy1 - ny0;
      rt[0] = prev[ly * nrw + lx];
      rt[1] = prev[ly * nrw + lx1];
      rt[2] = prev[ly1 * nrw + lx];
      rt[3] = prev[ly1 * nrw + lx1];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
        if (x < rx0 || x > rx1 || y < ry0 || y > ry1) continue;
        const size_t pix = ((size_t)b * d.h + y) * d.w + x;
        float4 Lv[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float4 g = ld_px<C>(d.g[i], pix);
          const float4 up = up_quad_lap<C>(e, gt[i]);
          Lv[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        }
        const float4 B = blend_any<C, N, KIND>(Lv, d.mask, (size_t)y * d.w + x);
        const float4 up = up_quad_rec<C>(e, rt);
        const float4 R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
        if (j == 0)
          st_px<C>(a.r_out, pix, R);
        else
          buf[cur][(y - ry0) * rw + (x - rx0)] = R;
      }
    }
  }
}

template <int C, int N, int KIND>
__global__ void __launch_bounds__(256) up_tail_kernel(UpTailArgs<N> a) {
  __shared__ float4 lds[kTailLdsFloat4];
  (void)a;
  up_tail_body<C, N, KIND>(kernel_arg<UpTailArgs<N>>(), hw_blk(), lds);
}

// The same tail with every level's footprint (Gaussians and mask) read into
// LDS first, in one burst: each thread issues all of its loads (for all
// levels) before the first LDS write, so the workgroup pays one memory round
// trip instead of one per level, and the level-by-level reconstruction then
// runs from LDS.  T x T base tiles; up to kTailLds levels above the base.
// Same helpers and operation order as up_tail_kernel: the same bits.
constexpr int kTailLds = 4;
constexpr int kTailLdsTile = 16;

// LDS of the burst tail, in float4: the footprints of every level (Gaussians,
// then the mask), R of two levels, and the footprint bounds.
template <int N, int KIND, int T>
__host__ __device__ constexpr int up_tail_lds_float4() {
  constexpr int M = KIND == kBlend2 ? 1 : N;
  constexpr int RU = T / 2 + 2, CAP = T * T + kTailLds * RU * RU;
  return N * CAP + 2 * RU * RU + (M * CAP + 3) / 4 + ((kTailLds + 1) * 4 + kTailLds + 2 + 3) / 4;
}

// Runs on 256 threads; lds: up_tail_lds_float4<N, KIND, T>() float4.
template <int C, int N, int KIND, int T>
__device__ __forceinline__ void up_tail_lds_body(const UpTailArgs<N>& a, Blk blk, float4* lds) {
  constexpr int M = KIND == kBlend2 ? 1 : N;
  constexpr int RU = T / 2 + 2;                      // footprint edge above the base
  constexpr int CAP = T * T + kTailLds * RU * RU;    // pixels of all footprints
  constexpr int SB = (T * T + 255) / 256;            // base-level pixels per thread
  constexpr int SU = (RU * RU + 255) / 256;          // pixels per thread at each level above
  float4* gl = lds;                                  // [level][image][footprint pixel]
  float4(*rb)[RU * RU] = (float4(*)[RU * RU])(lds + N * CAP);
  float* ml = (float*)(lds + N * CAP + 2 * RU * RU);  // [level][footprint pixel][M]
  int(*reg)[4] = (int(*)[4])(ml + ((M * CAP + 3) & ~3));  // footprint [x0, x1] x [y0, y1]
  int* cum = &reg[kTailLds + 1][0];                  // pixels of the footprints before level j
  const int t = blk.x;
  const int tx = t % a.tiles_x, ty = t / a.tiles_x;
  const int b = blk.z;
  const int tid = flat_tid();
  const int n = a.n;
  if (tid == 0) {
    int x0 = a.wx0 + tx * T, y0 = a.wy0 + ty * T;
    int x1 = min(x0 + T - 1, a.wx1), y1 = min(y0 + T - 1, a.wy1);
    reg[0][0] = x0; reg[0][1] = x1; reg[0][2] = y0; reg[0][3] = y1;
    cum[0] = 0;
    cum[1] = (x1 - x0 + 1) * (y1 - y0 + 1);
    for (int j = 1; j < n; ++j) {
      x0 >>= 1;
      y0 >>= 1;
      x1 = min((x1 >> 1) + 1, a.lv[j].w - 1);
      y1 = min((y1 >> 1) + 1, a.lv[j].h - 1);
      reg[j][0] = x0; reg[j][1] = x1; reg[j][2] = y0; reg[j][3] = y1;
      cum[j + 1] = cum[j] + (x1 - x0 + 1) * (y1 - y0 + 1);
    }
  }
  __syncthreads();
  // 1. the burst: every footprint pixel of every level.  The level loop is
  // unrolled (uniform level: descriptors by scalar loads); each thread holds
  // its pixels of all levels in registers before the first LDS write.
  float4 gv[SB + kTailLds * SU][N];
  float mv[SB + kTailLds * SU][M];
#pragma unroll
  for (int j = 0; j <= kTailLds; ++j) {
    if (j < n) {
      const TailLevel<N>& d = a.lv[j];
      const int rw = reg[j][1] - reg[j][0] + 1, area = cum[j + 1] - cum[j];
#pragma unroll
      for (int s = 0; s < (j == 0 ? SB : SU); ++s) {
        const int idx = s * 256 + tid;
        if (idx < area) {
          const int x = reg[j][0] + idx % rw, y = reg[j][2] + idx / rw;
          const size_t pix = ((size_t)b * d.h + y) * d.w + x;
          const int slot = (j == 0 ? 0 : SB + (j - 1) * SU) + s;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            // Three or more images are sparse: outside its valid region an
            // image's level is zero (see LevelDesc) and its buffer is not read.
            gv[slot][i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (KIND == kBlend2 || in_rect(d.z[i], x, y)) gv[slot][i] = ld_px<C>(d.g[i], pix);
          }
#pragma unroll
          for (int c = 0; c < M; ++c) mv[slot][c] = d.mask[((size_t)y * d.w + x) * M + c];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j <= kTailLds; ++j) {
    if (j < n) {
      const int area = cum[j + 1] - cum[j];
#pragma unroll
      for (int s = 0; s < (j == 0 ? SB : SU); ++s) {
        const int idx = s * 256 + tid;
        if (idx < area) {
          const int slot = (j == 0 ? 0 : SB + (j - 1) * SU) + s;
#pragma unroll
          for (int i = 0; i < N; ++i) gl[N * cum[j] + i * area + idx] = gv[slot][i];
#pragma unroll
          for (int c = 0; c < M; ++c) ml[(cum[j] + idx) * M + c] = mv[slot][c];
        }
      }
    }
  }
  __syncthreads();
  auto gpx = [&](int j, int i, int lidx) -> float4 { return gl[N * cum[j] + i * (cum[j + 1] - cum[j]) + lidx]; };
  int cur = 0;
  // 2. coarsest level: R = blend of the Gaussians
  {
    const int j = n - 1;
    const int rw = reg[j][1] - reg[j][0] + 1, area = cum[j + 1] - cum[j];
    (void)rw;
    for (int idx = tid; idx < area; idx += kWgThreads) {
      float4 Lv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) Lv[i] = gpx(j, i, idx);
      rb[cur][idx] = blend_any<C, N, KIND>(Lv, ml + cum[j] * M, (size_t)idx);
    }
  }
  // 3. levels n-2 .. 0 in quads, everything from LDS; the base writes R
  for (int j = n - 2; j >= 0; --j) {
    __syncthreads();
    cur ^= 1;
    const TailLevel<N>& d = a.lv[j];
    const TailLevel<N>& dn = a.lv[j + 1];
    const int rx0 = reg[j][0], rx1 = reg[j][1], ry0 = reg[j][2], ry1 = reg[j][3];
    const int rw = rx1 - rx0 + 1;
    const int qx0 = rx0 >> 1, qy0 = ry0 >> 1;
    const int nqx = (rx1 >> 1) - qx0 + 1, nqy = (ry1 >> 1) - qy0 + 1;
    const int nx0 = reg[j + 1][0], ny0 = reg[j + 1][2], nrw = reg[j + 1][1] - nx0 + 1;
    const float4* prev = rb[cur ^ 1];
    const float* mk = ml + cum[j] * M;
    for (int q = tid; q < nqx * nqy; q += kWgThreads) {
      const int qx = qx0 + q % nqx, qy = qy0 + q / nqx;
      const int qx1 = min(qx + 1, dn.w - 1), qy1 = min(qy + 1, dn.h - 1);
      const int lx = qx - nx0, lx1 = qx1 - nx0, ly = qy - ny0, ly1 = qy1 - ny0;
      float4 gt[N][4];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        gt[i][0] = gpx(j + 1, i, ly * nrw + lx);
        gt[i][1] = gpx(j + 1, i, ly * nrw + lx1);
        gt[i][2] = gpx(j + 1, i, ly1 * nrw + lx);
        gt[i][3] = gpx(j + 1, i, ly1 * nrw + lx1);
      }
      float4 rt[4];
      rt[0] = prev[ly * nrw + lx];
      rt[1] = prev[ly * nrw + lx1];
      rt[2] = prev[ly1 * nrw + lx];
      rt[3] = prev[ly1 * nrw + lx1];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
        if (x < rx0 || x > rx1 || y < ry0 || y > ry1) continue;
        const int lidx = (y - ry0) * rw + (x - rx0);
        float4 Lv[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const float4 g = gpx(j, i, lidx);
          const float4 up = up_quad_lap<C>(e, gt[i]);
          Lv[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        }
        const float4 B = blend_any<C, N, KIND>(Lv, mk, (size_t)lidx);
        const float4 up = up_quad_rec<C>(e, rt);
        const float4 R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
        if (j == 0)
          st_px<C>(a.r_out, ((size_t)b * d.h + y) * d.w + x, R);
        else
          rb[cur][lidx] = R;
      }
    }
  }
}

template <int C, int N, int KIND, int T>
__global__ void __launch_bounds__(256) up_tail_lds_kernel(UpTailArgs<N> a) {
  __shared__ float4 lds[up_tail_lds_float4<N, KIND, T>()];
  (void)a;
  up_tail_lds_body<C, N, KIND, T>(kernel_arg<UpTailArgs<N>>(), hw_blk(), lds);
}

// ---------------------------------------------------------------------------
// one level in 2x2 quads (fields named as level 0 and its next level 1)
// ---------------------------------------------------------------------------
template <typename T0, int N>
struct UpQuadArgs {
  const T0* g0[N];            // the level's images (tight, batch-major)
  const void* mask0;          // the level's mask (see Mask0Mode)
  const float* g1[N];         // next level's Gaussians
  const float* r1;            // next level's reconstruction, or nullptr: coarsest level, blend only
  int w0, h0, w1, h1;
  float* f32_out;             // float output (tight), or nullptr
  uint8_t* u8_out;            // uint8 canvas, or nullptr
  size_t u8_pitch;
  int u8_w, u8_h, ox, oy;
  int win_x0, win_y0, win_x1, win_y1;  // pixels produced: [x0, x1] x [y0, y1]
  int qx0, qy0, nqx, nqy;              // quads covering the window
  ImgRect z[N];                        // valid region of each image at this level (see LevelDesc)
  int tiles_x, tiles_y;                // 64 x 4 quads per workgroup
  const float* wn;                     // N-image blend: planar normalised weights, or nullptr
  size_t wn_plane;                     // floats per plane (w0 * h0)
};

// ---------------------------------------------------------------------------
// Three or more images: the images of a workgroup's tile
// ---------------------------------------------------------------------------
// A workgroup covers 64 x 4 quads (128 x 8 pixels).  Outside its valid region
// an image's Laplacian is zero, so only the images whose region meets the
// tile can contribute there.  Up to kQuadSlots of them are carried in slots,
// in index order; the slot count S is a template parameter (the kernel
// switches on the tile's count, a uniform value), so every slot is in use and
// all loads are issued without a branch between them.  The blend of the slots
// gives the bits of the blend of all images (blendN_weighted).  A tile met by
// more images takes the per-pixel path.
constexpr int kQuadSlots = 4;

// The images meeting the tile [x0, x1] x [y0, y1] (uniform): their number,
// their set (bit i: image i), and the first kQuadSlots of them in idx, in
// index order (0 past the end).
template <int N>
__device__ __forceinline__ int tile_images(const ImgRect* z, int x0, int y0, int x1, int y1, int* idx,
                                           unsigned* set) {
  unsigned act = 0;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (z[i].x0 <= x1 && z[i].x1 >= x0 && z[i].y0 <= y1 && z[i].y1 >= y0) act |= 1u << i;
  *set = act;
  const int n = __builtin_popcount(act);
#pragma unroll
  for (int s = 0; s < kQuadSlots; ++s) {
    idx[s] = act ? __builtin_ctz(act) : 0;
    act &= act - 1u;
  }
  return n;
}

// The quad's pixels: position, whether in the window, and the clamped pixel
// loaded (pixels outside the window load the window's first pixel).
struct QuadPix {
  int x[4], y[4];
  bool in[4];
  size_t pix[4];  // batched image index
  size_t mpx[4];  // unbatched (mask) index
};
template <typename Args>
__device__ __forceinline__ QuadPix quad_pixels(const Args& a, int qx, int qy, int b) {
  QuadPix q;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
    q.x[e] = x;
    q.y[e] = y;
    q.in[e] = x >= a.win_x0 && x <= a.win_x1 && y >= a.win_y0 && y <= a.win_y1;
    const int cx = q.in[e] ? x : a.win_x0, cy = q.in[e] ? y : a.win_y0;
    q.pix[e] = ((size_t)b * a.h0 + cy) * a.w0 + cx;
    q.mpx[e] = (size_t)cy * a.w0 + cx;
  }
  return q;
}

// R of level pixel (x, y) of batch b to the float output and/or the canvas.
template <int C, typename Args>
__device__ __forceinline__ void quad_store(const Args& a, int x, int y, int b, const float4& R) {
  if (a.f32_out) st_px<C>(a.f32_out, ((size_t)b * a.h0 + y) * a.w0 + x, R);
  if (a.u8_out) {
    const int X = x + a.ox, Y = y + a.oy;
    if (X < a.u8_w && Y < a.u8_h)
      st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
  }
}

// Level 0 of the N- and 3-image blends from the seam labels: the labelled
// image's Laplacian where it is opaque (3 channels: always), else the first
// slot of largest alpha (see up_quad_kernel).  Images outside the slots are
// zero at the pixel and never win the alpha search.
template <int C, int S>
__device__ __forceinline__ float4 label_blend(const float4* Lv, const int* idx, int lab) {
  float4 B = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool done = false;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (idx[s] == lab && (C == 3 || Lv[s].w != 0.0f)) {
      B = C == 4 ? Lv[s] : make_float4(Lv[s].x, Lv[s].y, Lv[s].z, 0.0f);
      done = true;
    }
  }
  if constexpr (C == 4) {
    if (!done) {
      float amax = 0.0f;
#pragma unroll
      for (int s = 0; s < S; ++s) {
        if (Lv[s].w > amax) {
          amax = Lv[s].w;
          B = Lv[s];
        }
      }
    }
  }
  return B;
}

// Slot of image j under the image set MASK (its rank among the set's images).
__host__ __device__ constexpr int slot_of(int mask, int j) { return __builtin_popcount(mask & ((1 << j) - 1)); }
// Image of slot s under the image set MASK.
__host__ __device__ constexpr int image_of(int mask, int s) {
  return (mask & 1) ? (s == 0 ? 0 : 1 + image_of(mask >> 1, s - 1)) : 1 + image_of(mask >> 1, s);
}

// The coarser level's taps of a 64 x 4 quad tile: quad (qx, qy) reads
// positions qx..qx+1 x qy..qy+1, so the tile reads (64 + 1) x (4 + 1) of
// them, each clamped to the level like a thread's own taps.  With a coarser
// level, the tile stages R's and each slot image's taps there in LDS (two
// loads per thread and image instead of four, and no tap registers held
// across the loads of the level's own pixels).
constexpr int kQuadTapW = 64 + 1, kQuadTapH = 4 + 1, kQuadTaps = kQuadTapW * kQuadTapH;
template <int N>
__host__ __device__ constexpr int quad_lds_float4() {
  return ((N < kQuadSlots ? N : kQuadSlots) + 1) * kQuadTaps;
}

// A uint8 pixel as loaded (4 bytes; 3 channels: packed), converted where it
// is used: the same values ld_px gives.
template <int C>
__device__ __forceinline__ uint32_t ld_raw_u8(const uint8_t* base, size_t pix) {
  if constexpr (C == 4) {
    return *reinterpret_cast<const uint32_t*>(base + pix * 4);
  } else {
    const uint8_t* p = base + pix * 3;
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
  }
}
template <int C>
__device__ __forceinline__ float4 cvt_raw_u8(uint32_t r) {
  return make_float4((float)(r & 0xffu), (float)((r >> 8) & 0xffu), (float)((r >> 16) & 0xffu),
                     C == 4 ? (float)(r >> 24) : 0.0f);
}

// The labelled slot of each pixel of a quad (-1: the label is not a slot)
// and that slot's level-0 pixel as loaded (slot 0's where there is none).
template <int C, int N, int S>
__device__ __forceinline__ void label_slot_pixels(const UpQuadArgs<uint8_t, N>& a, const int* idx, const QuadPix& q,
                                                  const int* lab, int* sl, uint32_t* rv) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    sl[e] = -1;
    const uint8_t* g = a.g0[idx[0]];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (idx[s] == lab[e]) {
        sl[e] = s;
        g = a.g0[idx[s]];
      }
    }
    rv[e] = ld_raw_u8<C>(g, q.pix[e]);
  }
}

// Level 0 of the label path from the tile's level-1 footprint in LDS (R at
// lds, slot s's Gaussian at lds + (s + 1) * kQuadTaps; NEXT).  The blend of
// a pixel is the labelled image's Laplacian where that image is opaque (3
// channels: always), else the first slot of largest alpha (label_blend):
// the labelled slot's pixel was loaded (sl, rv), the others are loaded only
// where that one is transparent or not a slot.  Every value is computed by
// the functions, on the inputs, of the general slot path.  The same steps as
// the end of up_quad_label_slots, which keeps its own inline copy: called
// from there, this form costs that kernel 4 more VGPRs and a wavefront per
// SIMD (84 against 80 with eight images).
template <int C, int N, bool NEXT, int S>
__device__ __forceinline__ void label_level0(const UpQuadArgs<uint8_t, N>& a, const int* idx, const QuadPix& q,
                                             const int* sl, const uint32_t* rv, const float4* lds, int b) {
  const int t0 = (int)threadIdx.y * kQuadTapW + (int)threadIdx.x;  // tap (qx, qy) in the footprint
  float4 rt[4];
  if (NEXT) {
    rt[0] = lds[t0];
    rt[1] = lds[t0 + 1];
    rt[2] = lds[t0 + kQuadTapW];
    rt[3] = lds[t0 + kQuadTapW + 1];
  }
  // Laplacian of slot s at pixel e of the quad (zero outside its region)
  auto lap_of = [&](int s, uint32_t v, const ImgRect& zs, int e) -> float4 {
    if (!in_rect(zs, q.x[e], q.y[e])) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 g = cvt_raw_u8<C>(v);
    if (!NEXT) return g;
    const float4* tp = lds + (s + 1) * kQuadTaps + t0;
    const float4 gt[4] = {tp[0], tp[1], tp[kQuadTapW], tp[kQuadTapW + 1]};
    const float4 up = up_quad_lap<C>(e, gt);  // e is compile-time after unrolling
    return make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
  };
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!q.in[e]) continue;
    ImgRect zs = {0, 0, -1, -1};
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (s == sl[e]) zs = a.z[idx[s]];
    float4 B = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool done = false;
    if (sl[e] >= 0) {
      const float4 L = lap_of(sl[e], rv[e], zs, e);
      if (C == 3 || L.w != 0.0f) {
        B = C == 4 ? L : make_float4(L.x, L.y, L.z, 0.0f);
        done = true;
      }
    }
    if constexpr (C == 4) {
      if (!done) {  // alpha fallback: the first slot of largest alpha
        float amax = 0.0f;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float4 L = lap_of(s, ld_raw_u8<C>(a.g0[idx[s]], q.pix[e]), a.z[idx[s]], e);
          if (L.w > amax) {
            amax = L.w;
            B = L;
          }
        }
      }
    }
    float4 R = B;
    if (NEXT) {
      const float4 up = up_quad_rec<C>(e, rt);
      R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    }
    quad_store<C>(a, q.x[e], q.y[e], b, R);
  }
}

// Level 0 of the N-image blend from the seam labels, through the tile's S
// slot images (uint8 level 0).  The blend of a pixel is the labelled image's
// Laplacian where that image is opaque (3 channels: always), else the first
// slot of largest alpha (label_blend), so each pixel loads and computes the
// labelled slot's pixel only (the labels are read first), and all slots'
// only where that one is transparent or not a slot.  Every value is computed
// by the functions, on the inputs, of the general slot path.
template <int C, int N, bool NEXT, int S>
__device__ __forceinline__ void up_quad_label_slots(const UpQuadArgs<uint8_t, N>& a, const int* idx_rt, int tqx,
                                                    int tqy, int b, float4* lds) {
  constexpr int SA = S > 0 ? S : 1;  // array extent
  int idx[SA];
#pragma unroll
  for (int s = 0; s < S; ++s) idx[s] = idx_rt[s];
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const int qx = tqx + lx, qy = tqy + ly;
  const QuadPix q = quad_pixels(a, qx, qy, b);
  int lab[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) lab[e] = static_cast<const uint8_t*>(a.mask0)[q.mpx[e]];
  // the tile's taps of R and of each slot image, staged in LDS first
  constexpr int kRounds = (kQuadTaps + kWgThreads - 1) / kWgThreads;
  const int tid = ly * 64 + lx;
  if (NEXT) {
    size_t tap[kRounds];
    float4 st[S + 1][kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = min(tid + r * kWgThreads, kQuadTaps - 1);
      const int px = min(tqx + p % kQuadTapW, a.w1 - 1), py = min(tqy + p / kQuadTapW, a.h1 - 1);
      tap[r] = ((size_t)b * a.h1 + py) * a.w1 + px;
      st[0][r] = ld_px<C>(a.r1, tap[r]);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float* g1 = a.g1[idx[s]];
#pragma unroll
      for (int r = 0; r < kRounds; ++r) st[s + 1][r] = ld_px<C>(g1, tap[r]);
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = tid + r * kWgThreads;
      if (p < kQuadTaps) {
#pragma unroll
        for (int s = 0; s <= S; ++s) lds[s * kQuadTaps + p] = st[s][r];
      }
    }
  }
  // each pixel's labelled slot (-1: none) and its pixel as loaded
  int sl[4];
  uint32_t rv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    sl[e] = -1;
    const uint8_t* g = a.g0[idx[0]];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (idx[s] == lab[e]) {
        sl[e] = s;
        g = a.g0[idx[s]];
      }
    }
    rv[e] = ld_raw_u8<C>(g, q.pix[e]);
  }
  const int t0 = ly * kQuadTapW + lx;  // tap (qx, qy) in the staged footprint
  if (NEXT) __syncthreads();
  float4 rt[4];
  if (NEXT) {
    rt[0] = lds[t0];
    rt[1] = lds[t0 + 1];
    rt[2] = lds[t0 + kQuadTapW];
    rt[3] = lds[t0 + kQuadTapW + 1];
  }
  // Laplacian of slot s at pixel e of the quad (zero outside its region)
  auto lap_of = [&](int s, uint32_t v, const ImgRect& zs, int e) -> float4 {
    if (!in_rect(zs, q.x[e], q.y[e])) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 g = cvt_raw_u8<C>(v);
    if (!NEXT) return g;
    const float4* tp = lds + (s + 1) * kQuadTaps + t0;
    const float4 gt[4] = {tp[0], tp[1], tp[kQuadTapW], tp[kQuadTapW + 1]};
    const float4 up = up_quad_lap<C>(e, gt);  // e is compile-time after unrolling
    return make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
  };
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!q.in[e]) continue;
    ImgRect zs = {0, 0, -1, -1};
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (s == sl[e]) zs = a.z[idx[s]];
    float4 B = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    bool done = false;
    if (sl[e] >= 0) {
      const float4 L = lap_of(sl[e], rv[e], zs, e);
      if (C == 3 || L.w != 0.0f) {
        B = C == 4 ? L : make_float4(L.x, L.y, L.z, 0.0f);
        done = true;
      }
    }
    if constexpr (C == 4) {
      if (!done) {  // alpha fallback: the first slot of largest alpha
        float amax = 0.0f;
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const float4 L = lap_of(s, ld_raw_u8<C>(a.g0[idx[s]], q.pix[e]), a.z[idx[s]], e);
          if (L.w > amax) {
            amax = L.w;
            B = L;
          }
        }
      }
    }
    float4 R = B;
    if (NEXT) {
      const float4 up = up_quad_rec<C>(e, rt);
      R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    }
    quad_store<C>(a, q.x[e], q.y[e], b, R);
  }
}

// One quad of a level of the N- or 3-image blend through the tile's S slot
// images: idx[0..S-1], or, when MASK >= 0, the images of the set MASK (a
// compile-time set: the 3-image blend places each image's Laplacian at its
// own position).  Weights: the seam labels (level 0), the planar normalised
// weights (N-image levels >= 1) or the interleaved 3-image mask.  (tqx, tqy):
// the tile's first quad; every thread of the workgroup calls this (the tap
// staging has a barrier); those past the window store nothing (none of their
// pixels is inside it).
template <typename T0, int C, int N, int KIND, int MODE, bool NEXT, int S, int MASK = -1>
__device__ __forceinline__ void up_quad_slots(const UpQuadArgs<T0, N>& a, const int* idx_rt, int tqx, int tqy,
                                              int b, float4* lds) {
  if constexpr (MODE == kMask0Label && MASK < 0 && std::is_same<T0, uint8_t>::value) {
    up_quad_label_slots<C, N, NEXT, S>(a, idx_rt, tqx, tqy, b, lds);
    return;
  }
  constexpr int SA = S > 0 ? S : 1;  // array extent
  int idx[SA];
#pragma unroll
  for (int s = 0; s < S; ++s) idx[s] = MASK >= 0 ? image_of(MASK, s) : idx_rt[s];
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const int qx = tqx + lx, qy = tqy + ly;
  const QuadPix q = quad_pixels(a, qx, qy, b);
  // the tile's taps of R and of each slot image first, stored in LDS before
  // the slots' pixels and the weights are loaded (no tap registers live
  // across those loads)
  constexpr int kRounds = (kQuadTaps + kWgThreads - 1) / kWgThreads;
  const int tid = ly * 64 + lx;
  size_t tap[kRounds];
  float4 st[S + 1][kRounds];
  if (NEXT) {
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = min(tid + r * kWgThreads, kQuadTaps - 1);
      const int px = min(tqx + p % kQuadTapW, a.w1 - 1), py = min(tqy + p / kQuadTapW, a.h1 - 1);
      tap[r] = ((size_t)b * a.h1 + py) * a.w1 + px;
      st[0][r] = ld_px<C>(a.r1, tap[r]);
    }
  }
  ImgRect z[SA];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    z[s] = a.z[idx[s]];
    if (NEXT) {
      const float* g1 = a.g1[idx[s]];
#pragma unroll
      for (int r = 0; r < kRounds; ++r) st[s + 1][r] = ld_px<C>(g1, tap[r]);
    }
  }
  if (NEXT) {
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = tid + r * kWgThreads;
      if (p < kQuadTaps) {
#pragma unroll
        for (int s = 0; s <= S; ++s) lds[s * kQuadTaps + p] = st[s][r];
      }
    }
  }
  // The N-image blend loads each slot's pixels inside its valid region only
  // (elsewhere the Laplacian is zero and the pixel unused; lanes outside load
  // nothing): its images are sparse (8 x 3840x2160 -> 14000x2600: 146 us
  // against 153 us for levels 1..5).  The 3-image class's images cover the
  // canvas and lose by the branches (67 us against 61 us at 6000x2000).
  constexpr bool kSparseLoads = KIND == kBlendN;
  float4 gv[SA][4];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const T0* g0 = a.g0[idx[s]];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      gv[s][e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (!kSparseLoads || in_rect(z[s], q.x[e], q.y[e])) gv[s][e] = ld_px<C>(g0, q.pix[e]);
    }
  }
  int lab[4];
  float wv[SA][4];
  float mk[4][3];
  if constexpr (MODE == kMask0Label) {
#pragma unroll
    for (int e = 0; e < 4; ++e) lab[e] = static_cast<const uint8_t*>(a.mask0)[q.mpx[e]];
  } else if constexpr (KIND == kBlend3) {
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int c = 0; c < 3; ++c) mk[e][c] = static_cast<const float*>(a.mask0)[q.mpx[e] * 3 + c];
  } else {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      // Outside the slot's valid region its Laplacian is zero: its weight is
      // dropped (RGBA) or multiplies zeros (RGB), so it is not loaded there.
      const float* w = a.wn + (size_t)idx[s] * a.wn_plane;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        wv[s][e] = 0.0f;
        if (!kSparseLoads || in_rect(z[s], q.x[e], q.y[e])) wv[s][e] = w[q.mpx[e]];
      }
    }
  }
  const int t0 = ly * kQuadTapW + lx;  // tap (qx, qy) in the staged footprint
  if (NEXT) __syncthreads();
  // each slot's Laplacian at the quad's pixels (zero outside its region)
  float4 Lq[SA][4];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    float4 gt[4];
    if (NEXT) {
      const float4* tp = lds + (s + 1) * kQuadTaps + t0;
      gt[0] = tp[0];
      gt[1] = tp[1];
      gt[2] = tp[kQuadTapW];
      gt[3] = tp[kQuadTapW + 1];
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      Lq[s][e] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (in_rect(z[s], q.x[e], q.y[e])) {
        const float4 g = gv[s][e];
        if (NEXT) {
          const float4 up = up_quad_lap<C>(e, gt);
          Lq[s][e] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        } else {
          Lq[s][e] = g;
        }
      }
    }
  }
  float4 rt[4];
  if (NEXT) {
    rt[0] = lds[t0];
    rt[1] = lds[t0 + 1];
    rt[2] = lds[t0 + kQuadTapW];
    rt[3] = lds[t0 + kQuadTapW + 1];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!q.in[e]) continue;
    float4 Lv[SA];
#pragma unroll
    for (int s = 0; s < S; ++s) Lv[s] = Lq[s][e];
    float4 B;
    if constexpr (MODE == kMask0Label) {
      B = label_blend<C, S>(Lv, idx, lab[e]);
    } else if constexpr (KIND == kBlend3) {
      static_assert(MASK >= 0, "the 3-image blend takes a compile-time image set");
      float4 L3[3];  // the three images' Laplacians, zero for those outside the set
#pragma unroll
      for (int j = 0; j < 3; ++j)
        L3[j] = ((MASK >> j) & 1) ? Lv[slot_of(MASK, j)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      B = blend3_px<C>(L3, mk[e]);
    } else {
      float w[SA];
#pragma unroll
      for (int s = 0; s < S; ++s) w[s] = wv[s][e];
      B = blendN_weighted<C, S>(Lv, w);
    }
    float4 R = B;
    if (NEXT) {
      const float4 up = up_quad_rec<C>(e, rt);
      R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    }
    quad_store<C>(a, q.x[e], q.y[e], b, R);
  }
}

// LDS of an up_quad launch: the tap footprint where the slot path or the
// dense path (one or two images) runs with a coarser level, else none.
template <int N, int KIND, int MODE, bool NEXT>
__host__ __device__ constexpr int up_quad_lds_float4() {
  return ((N > 2 && NEXT) || (N <= 2 && NEXT && MODE != kMask0Label)) ? quad_lds_float4<N>() : 1;
}

// NEXT: the level has a coarser level (false at the coarsest: blend only).
// A template parameter rather than a branch, so that the taps of the coarser
// level and the level's own pixels are loaded in one straight run.
// Runs on 64 x 4 threads; lds: up_quad_lds_float4<N, KIND, MODE, NEXT>()
// float4 for the staged taps.
template <typename T0, int C, int N, int KIND, int MODE, bool NEXT>
__device__ __forceinline__ void up_quad_body(const UpQuadArgs<T0, N>& a, Blk blk, float4* lds) {
  constexpr bool has_next = NEXT;
  const int t = blk.x;
  const int tqx = a.qx0 + (t % a.tiles_x) * 64, tqy = a.qy0 + (t / a.tiles_x) * 4;
  const int qx = tqx + (int)threadIdx.x;
  const int qy = tqy + (int)threadIdx.y;
  const int b = blk.z;
  // Three or more images: the tile's images in slots when they fit and the
  // weights allow.  At level 0 (labels) the slot path computes only the
  // labelled image's Laplacian per pixel (up_quad_label_slots); for the
  // 3-image class, whose images cover the whole canvas, it beats the
  // per-pixel path below, which loads only the labelled image but only after
  // its label (59 against 71 us at 6000x2000).
  if constexpr (N > 2) {
    if (MODE == kMask0Label || KIND == kBlend3 || a.wn != nullptr) {
      int idx[kQuadSlots];
      unsigned set = 0;
      const int n = tile_images<N>(a.z, 2 * tqx, 2 * tqy, 2 * tqx + 127, 2 * tqy + 7, idx, &set);
      if constexpr (KIND == kBlend3 && MODE != kMask0Label) {
        switch (set) {  // uniform
          case 0: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 0, 0>(a, idx, tqx, tqy, b, lds); return;
          case 1: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 1, 1>(a, idx, tqx, tqy, b, lds); return;
          case 2: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 1, 2>(a, idx, tqx, tqy, b, lds); return;
          case 3: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 2, 3>(a, idx, tqx, tqy, b, lds); return;
          case 4: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 1, 4>(a, idx, tqx, tqy, b, lds); return;
          case 5: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 2, 5>(a, idx, tqx, tqy, b, lds); return;
          case 6: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 2, 6>(a, idx, tqx, tqy, b, lds); return;
          default: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 3, 7>(a, idx, tqx, tqy, b, lds); return;
        }
      } else {
        switch (n) {  // uniform
          case 0: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 0>(a, idx, tqx, tqy, b, lds); return;
          case 1: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 1>(a, idx, tqx, tqy, b, lds); return;
          case 2: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 2>(a, idx, tqx, tqy, b, lds); return;
          case 3: up_quad_slots<T0, C, N, KIND, MODE, NEXT, 3>(a, idx, tqx, tqy, b, lds); return;
          case 4:
            if constexpr (N >= 4) {
              up_quad_slots<T0, C, N, KIND, MODE, NEXT, 4>(a, idx, tqx, tqy, b, lds);
              return;
            }
            break;
          default: break;  // more images than slots: the per-pixel path below
        }
      }
    }
  }
  // The dense path with a coarser level stages the tile's taps with every
  // thread (a barrier); every other path drops the threads past the window.
  constexpr bool kDenseStaged = N <= 2 && NEXT && MODE != kMask0Label;
  const bool active = qx < a.qx0 + a.nqx && qy < a.qy0 + a.nqy;
  if (!kDenseStaged && !active) return;
  if constexpr (MODE == kMask0Label) {
    // Level 0 of the N- and 3-image blends on a tile met by more images
    // than there are slots.  The mask is one-hot, so the
    // blend (blendN_px / blend3_px with one weight 1 and the others 0, whose
    // terms are all +-0) is the labelled image's Laplacian where that image
    // is opaque, and elsewhere the alpha fallback: the first image of largest
    // alpha (3 channels: the labelled image, or zero).  Only the images the
    // result needs are loaded.
    float4 rt[4];
    const int qx1 = has_next ? min(qx + 1, a.w1 - 1) : 0, qy1 = has_next ? min(qy + 1, a.h1 - 1) : 0;
    const size_t l0 = ((size_t)b * a.h1 + qy) * a.w1, l1 = ((size_t)b * a.h1 + qy1) * a.w1;
    if (has_next) {
      rt[0] = ld_px<C>(a.r1, l0 + qx);
      rt[1] = ld_px<C>(a.r1, l0 + qx1);
      rt[2] = ld_px<C>(a.r1, l1 + qx);
      rt[3] = ld_px<C>(a.r1, l1 + qx1);
    }
    const uint8_t* labels = static_cast<const uint8_t*>(a.mask0);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
      if (x < a.win_x0 || x > a.win_x1 || y < a.win_y0 || y > a.win_y1) continue;
      const size_t pix = ((size_t)b * a.h0 + y) * a.w0 + x;
      const int lab = labels[(size_t)y * a.w0 + x];
      // Laplacian of image i at this pixel (zero outside its valid region)
      auto lap = [&](int i, const T0* g0i, const float* g1i, const ImgRect& zi) -> float4 {
        (void)i;
        if (!in_rect(zi, x, y)) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 g = ld_px<C>(g0i, pix);
        if (!has_next) return g;
        float4 t4[4];
        t4[0] = ld_px<C>(g1i, l0 + qx);
        t4[1] = ld_px<C>(g1i, l0 + qx1);
        t4[2] = ld_px<C>(g1i, l1 + qx);
        t4[3] = ld_px<C>(g1i, l1 + qx1);
        const float4 up = up_quad_lap<C>(e, t4);
        return make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
      };
      float4 B = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      bool done = false;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        if (i != lab) continue;
        const float4 Li = lap(i, a.g0[i], a.g1[i], a.z[i]);
        if (C == 3 || Li.w != 0.0f) {
          B = C == 4 ? Li : make_float4(Li.x, Li.y, Li.z, 0.0f);
          done = true;
        }
      }
      if constexpr (C == 4) {
        if (!done) {
          float amax = 0.0f;
#pragma unroll
          for (int i = 0; i < N; ++i) {
            const float4 Li = lap(i, a.g0[i], a.g1[i], a.z[i]);
            if (Li.w > amax) {
              amax = Li.w;
              B = Li;
            }
          }
        }
      }
      float4 R = B;
      if (has_next) {
        const float4 up = up_quad_rec<C>(e, rt);
        R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
      }
      if (a.f32_out) st_px<C>(a.f32_out, pix, R);
      if (a.u8_out) {
        const int X = x + a.ox, Y = y + a.oy;
        if (X < a.u8_w && Y < a.u8_h)
          st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
      }
    }
    return;
  }
  if constexpr (N > 2) {
    // Three or more images on a tile met by more images than there are
    // slots, or without planar weights: one pixel of the quad at a time, each
    // image's taps loaded where needed (cache hits across the quad), so
    // registers hold one pixel's Laplacians rather than the quad's.
    float4 rt[4];
    const int qx1 = has_next ? min(qx + 1, a.w1 - 1) : 0, qy1 = has_next ? min(qy + 1, a.h1 - 1) : 0;
    const size_t l0 = ((size_t)b * a.h1 + qy) * a.w1, l1 = ((size_t)b * a.h1 + qy1) * a.w1;
    if (has_next) {
      rt[0] = ld_px<C>(a.r1, l0 + qx);
      rt[1] = ld_px<C>(a.r1, l0 + qx1);
      rt[2] = ld_px<C>(a.r1, l1 + qx);
      rt[3] = ld_px<C>(a.r1, l1 + qx1);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
      if (x < a.win_x0 || x > a.win_x1 || y < a.win_y0 || y > a.win_y1) continue;
      const size_t pix = ((size_t)b * a.h0 + y) * a.w0 + x;
      float4 Lv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) {
        Lv[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (!in_rect(a.z[i], x, y)) continue;
        const float4 g = ld_px<C>(a.g0[i], pix);
        if (has_next) {
          float4 t4[4];
          t4[0] = ld_px<C>(a.g1[i], l0 + qx);
          t4[1] = ld_px<C>(a.g1[i], l0 + qx1);
          t4[2] = ld_px<C>(a.g1[i], l1 + qx);
          t4[3] = ld_px<C>(a.g1[i], l1 + qx1);
          const float4 up = up_quad_lap<C>(e, t4);
          Lv[i] = make_float4(g.x - up.x, g.y - up.y, g.z - up.z, g.w);
        } else {
          Lv[i] = g;
        }
      }
      const float4 B = blend_level0<C, N, KIND, MODE>(Lv, a.mask0, (size_t)y * a.w0 + x);
      float4 R = B;
      if (has_next) {
        const float4 up = up_quad_rec<C>(e, rt);
        R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
      }
      if (a.f32_out) st_px<C>(a.f32_out, pix, R);
      if (a.u8_out) {
        const int X = x + a.ox, Y = y + a.oy;
        if (X < a.u8_w && Y < a.u8_h)
          st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
      }
    }
    return;
  }
  // One or two images (dense): the quad's loads all issued before any
  // arithmetic.  With a coarser level, the tile's taps of R and of each
  // image (the four taps each quad's pixels share) are staged in LDS, as in
  // the slot path.
  constexpr int kRounds = (kQuadTaps + kWgThreads - 1) / kWgThreads;
  const int tid = (int)threadIdx.y * 64 + (int)threadIdx.x;
  float4 st[N + 1][kRounds];
  if (has_next) {
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = min(tid + r * kWgThreads, kQuadTaps - 1);
      const int px = min(tqx + p % kQuadTapW, a.w1 - 1), py = min(tqy + p / kQuadTapW, a.h1 - 1);
      const size_t tap = ((size_t)b * a.h1 + py) * a.w1 + px;
      st[0][r] = ld_px<C>(a.r1, tap);
#pragma unroll
      for (int i = 0; i < N; ++i) st[i + 1][r] = ld_px<C>(a.g1[i], tap);
    }
  }
  // the level's pixels and mask values of the quad, all loaded before any
  // arithmetic (pixels outside the window load the window's first pixel)
  constexpr int M = KIND == kBlend2 ? 1 : N;
  float4 g[4][N];
  float mk[4][M];
  bool in[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
    in[e] = x >= a.win_x0 && x <= a.win_x1 && y >= a.win_y0 && y <= a.win_y1;
    const int cx = in[e] ? x : a.win_x0, cy = in[e] ? y : a.win_y0;
    const size_t pix = ((size_t)b * a.h0 + cy) * a.w0 + cx;
#pragma unroll
    for (int i = 0; i < N; ++i) g[e][i] = ld_px<C>(a.g0[i], pix);
    const size_t mp = (size_t)cy * a.w0 + cx;
    if constexpr (MODE == kMask0U8) {
      mk[e][0] = (float)static_cast<const uint8_t*>(a.mask0)[mp];
    } else {
#pragma unroll
      for (int c = 0; c < M; ++c) mk[e][c] = static_cast<const float*>(a.mask0)[mp * M + c];
    }
  }
  float4 gt[N][4];
  float4 rt[4];
  if (has_next) {
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
      const int p = tid + r * kWgThreads;
      if (p < kQuadTaps) {
#pragma unroll
        for (int i = 0; i <= N; ++i) lds[i * kQuadTaps + p] = st[i][r];
      }
    }
    __syncthreads();
    if (!active) return;  // past the window: nothing to store
    const int t0 = (int)threadIdx.y * kQuadTapW + (int)threadIdx.x;  // tap (qx, qy) in the staged footprint
#pragma unroll
    for (int i = 0; i <= N; ++i) {
      const float4* tp = lds + i * kQuadTaps + t0;
      float4* dst = i == 0 ? rt : gt[i - 1];
      dst[0] = tp[0];
      dst[1] = tp[1];
      dst[2] = tp[kQuadTapW];
      dst[3] = tp[kQuadTapW + 1];
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!in[e]) continue;
    const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
    float4 Lv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (has_next) {
        const float4 up = up_quad_lap<C>(e, gt[i]);
        Lv[i] = make_float4(g[e][i].x - up.x, g[e][i].y - up.y, g[e][i].z - up.z, g[e][i].w);
      } else {
        Lv[i] = g[e][i];
      }
    }
    const float4 B = blend_any<C, N, KIND>(Lv, mk[e], 0);  // the preloaded mask values
    float4 R = B;
    if (has_next) {
      const float4 up = up_quad_rec<C>(e, rt);
      R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    }
    if (a.f32_out) st_px<C>(a.f32_out, ((size_t)b * a.h0 + y) * a.w0 + x, R);
    if (a.u8_out) {
      const int X = x + a.ox, Y = y + a.oy;
      if (X < a.u8_w && Y < a.u8_h)
        st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
    }
  }
}

template <typename T0, int C, int N, int KIND, int MODE, bool NEXT>
__global__ void __launch_bounds__(256) up_quad_kernel(UpQuadArgs<T0, N> a) {
  __shared__ float4 lds[up_quad_lds_float4<N, KIND, MODE, NEXT>()];
  (void)a;
  up_quad_body<T0, C, N, KIND, MODE, NEXT>(kernel_arg<UpQuadArgs<T0, N>>(), hw_blk(), lds);
}

// ---------------------------------------------------------------------------
// two images: levels 1 and 0 of the up pass in one launch
// ---------------------------------------------------------------------------
// A level-0 workgroup (64 x 4 quads) reads R1 and G1 only on the
// (64 + 1) x (4 + 1) level-1 footprint of its tile.  Here it computes R1 on
// that footprint itself -- the Laplacians of level 1, their blend and the
// reconstruction from level 2, with level 2's (32 + 2) x (2 + 2) footprint
// staged in LDS -- and keeps it in LDS, so R1 never goes to memory and the
// level-1 launch disappears.  Every R1 value is the level-1 launch's (the same
// functions on the same taps: the footprint positions are clamped to the level
// as that launch's staged taps are).  Footprint positions outside the level-1
// window use level-2 data outside the level-2 window (not written this
// frame); no level-0 pixel of the window reads them.
constexpr int kL2TapW = 32 + 2, kL2TapH = 2 + 2, kL2Taps = kL2TapW * kL2TapH;

template <typename T0>
struct UpL10Args {
  UpQuadArgs<T0, 2> l0;  // level 0 (g1: level 1's Gaussians; r1 is not read)
  const float* mask1;    // level-1 mask
  const float* g2[2];    // level-2 Gaussians (NEXT1)
  const float* r2;       // level-2 reconstruction (NEXT1)
  int w2, h2;
};

__host__ __device__ constexpr int up_l10_lds_float4() { return 3 * kQuadTaps + 3 * kL2Taps; }

// NEXT1: level 1 has a coarser level (false: levels == 2, level 1 is blended only).
// Runs on 64 x 4 threads; lds: up_l10_lds_float4() float4.
template <typename T0, int C, int MODE, bool NEXT1>
__device__ __forceinline__ void up_l10_body(const UpL10Args<T0>& A, Blk blk, float4* lds) {
  const UpQuadArgs<T0, 2>& a = A.l0;
  const int t = blk.x;
  const int tqx = a.qx0 + (t % a.tiles_x) * 64, tqy = a.qy0 + (t / a.tiles_x) * 4;
  const int qx = tqx + (int)threadIdx.x;
  const int qy = tqy + (int)threadIdx.y;
  const int b = blk.z;
  const bool active = qx < a.qx0 + a.nqx && qy < a.qy0 + a.nqy;
  const int tid = (int)threadIdx.y * 64 + (int)threadIdx.x;
  constexpr int kRounds = (kQuadTaps + kWgThreads - 1) / kWgThreads;
  float4* lr1 = lds;                  // R1 on the level-1 footprint
  float4* lg1 = lds + kQuadTaps;      // G1 of image i at lg1 + i * kQuadTaps
  float4* l2 = lds + 3 * kQuadTaps;   // level 2: R2, then G2 of each image
  const int bx2 = tqx >> 1, by2 = tqy >> 1;
  // Every load of the workgroup first: level 2's footprint, level 1's
  // footprint (both Gaussians and the mask), the quad's level-0 pixels and mask.
  float4 s2[3];
  if constexpr (NEXT1) {
    const int p = min(tid, kL2Taps - 1);
    const int px = min(bx2 + p % kL2TapW, A.w2 - 1), py = min(by2 + p / kL2TapW, A.h2 - 1);
    const size_t tap = ((size_t)b * A.h2 + py) * A.w2 + px;
    s2[0] = ld_px<C>(A.r2, tap);
    s2[1] = ld_px<C>(A.g2[0], tap);
    s2[2] = ld_px<C>(A.g2[1], tap);
  }
  float4 s1[kRounds][2];
  float m1[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = min(tid + r * kWgThreads, kQuadTaps - 1);
    const int px = min(tqx + p % kQuadTapW, a.w1 - 1), py = min(tqy + p / kQuadTapW, a.h1 - 1);
    const size_t tap = ((size_t)b * a.h1 + py) * a.w1 + px;
    s1[r][0] = ld_px<C>(a.g1[0], tap);
    s1[r][1] = ld_px<C>(a.g1[1], tap);
    m1[r] = A.mask1[(size_t)py * a.w1 + px];
  }
  float4 g[4][2];
  float mk[4];
  bool in[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
    in[e] = x >= a.win_x0 && x <= a.win_x1 && y >= a.win_y0 && y <= a.win_y1;
    const int cx = in[e] ? x : a.win_x0, cy = in[e] ? y : a.win_y0;
    const size_t pix = ((size_t)b * a.h0 + cy) * a.w0 + cx;
    g[e][0] = ld_px<C>(a.g0[0], pix);
    g[e][1] = ld_px<C>(a.g0[1], pix);
    const size_t mp = (size_t)cy * a.w0 + cx;
    if constexpr (MODE == kMask0U8)
      mk[e] = (float)static_cast<const uint8_t*>(a.mask0)[mp];
    else
      mk[e] = static_cast<const float*>(a.mask0)[mp];
  }
  if constexpr (NEXT1) {
    if (tid < kL2Taps) {
#pragma unroll
      for (int k = 0; k < 3; ++k) l2[k * kL2Taps + tid] = s2[k];
    }
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = tid + r * kWgThreads;
    if (p < kQuadTaps) {
      lg1[p] = s1[r][0];
      lg1[kQuadTaps + p] = s1[r][1];
    }
  }
  __syncthreads();
  // R1 on the footprint (the level-1 launch's quad arithmetic at each position)
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = tid + r * kWgThreads;
    if (p >= kQuadTaps) continue;
    const int x1 = min(tqx + p % kQuadTapW, a.w1 - 1), y1 = min(tqy + p / kQuadTapW, a.h1 - 1);
    const int e1 = (x1 & 1) | ((y1 & 1) << 1);
    float4 Lv[2] = {s1[r][0], s1[r][1]};
    float4 rt[4];
    if constexpr (NEXT1) {
      const int t00 = ((y1 >> 1) - by2) * kL2TapW + ((x1 >> 1) - bx2);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const float4* gp = l2 + (i + 1) * kL2Taps + t00;
        const float4 gt[4] = {gp[0], gp[1], gp[kL2TapW], gp[kL2TapW + 1]};
        const float4 up = up_quad_lap<C>(e1, gt);
        Lv[i] = make_float4(Lv[i].x - up.x, Lv[i].y - up.y, Lv[i].z - up.z, Lv[i].w);
      }
      const float4* rp = l2 + t00;
      rt[0] = rp[0];
      rt[1] = rp[1];
      rt[2] = rp[kL2TapW];
      rt[3] = rp[kL2TapW + 1];
    }
    const float4 B = blend2_px<C>(Lv[0], Lv[1], m1[r]);
    float4 R = B;
    if constexpr (NEXT1) {
      const float4 up = up_quad_rec<C>(e1, rt);
      R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    }
    lr1[p] = R;
  }
  __syncthreads();
  if (!active) return;  // past the window: nothing to store
  // Level 0 from the footprint in LDS, as the dense staged path of up_quad_body.
  const int t0 = (int)threadIdx.y * kQuadTapW + (int)threadIdx.x;  // tap (qx, qy) in the footprint
  float4 rt[4], gt[2][4];
#pragma unroll
  for (int i = 0; i <= 2; ++i) {
    const float4* tp = lds + i * kQuadTaps + t0;
    float4* dst = i == 0 ? rt : gt[i - 1];
    dst[0] = tp[0];
    dst[1] = tp[1];
    dst[2] = tp[kQuadTapW];
    dst[3] = tp[kQuadTapW + 1];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (!in[e]) continue;
    const int x = 2 * qx + (e & 1), y = 2 * qy + (e >> 1);
    float4 Lv[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const float4 up = up_quad_lap<C>(e, gt[i]);
      Lv[i] = make_float4(g[e][i].x - up.x, g[e][i].y - up.y, g[e][i].z - up.z, g[e][i].w);
    }
    const float4 B = blend2_px<C>(Lv[0], Lv[1], mk[e]);
    const float4 up = up_quad_rec<C>(e, rt);
    const float4 R = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
    if (a.f32_out) st_px<C>(a.f32_out, ((size_t)b * a.h0 + y) * a.w0 + x, R);
    if (a.u8_out) {
      const int X = x + a.ox, Y = y + a.oy;
      if (X < a.u8_w && Y < a.u8_h)
        st_px_u8<C>(a.u8_out + ((size_t)b * a.u8_h + Y) * a.u8_pitch + (size_t)X * C, R);
    }
  }
}

template <typename T0, int C, int MODE, bool NEXT1>
__global__ void __launch_bounds__(256) up_l10_kernel(UpL10Args<T0> a) {
  __shared__ float4 lds[up_l10_lds_float4()];
  (void)a;
  up_l10_body<T0, C, MODE, NEXT1>(kernel_arg<UpL10Args<T0>>(), hw_blk(), lds);
}

// ---------------------------------------------------------------------------
// three or more images: levels 1 and 0 of the up pass in one launch
// ---------------------------------------------------------------------------
// The N-image blend with uint8 level 0 and seam labels, as up_l10_body for
// two images: a level-0 workgroup computes R1 on its tile's (64 + 1) x
// (4 + 1) level-1 footprint itself, from level 2's (32 + 2) x (2 + 2)
// footprint staged in LDS, and keeps it in LDS, so R1 never goes to memory
// and the level-1 launch disappears.  The G1 footprint staged for R1 is the
// one level 0's Laplacians read.  Images go through the tile's slots: those
// whose level-0 region meets the tile or whose level-1 region meets the
// footprint, in index order -- a superset of the images that can be non-zero
// at any pixel either level computes, so the blends are the blends over all
// images (blendN_weighted, label_blend).  The host launches it only when no
// tile of the window has more than kQuadSlots of them.  Every R1 value is the
// level-1 launch's (the same functions on the same taps, footprint positions
// clamped to the level as that launch's taps are); footprint positions
// outside the level-1 window use level-2 data outside the level-2 window (not
// written this frame), which no level-0 pixel of the window reads.
template <int N>
struct UpL10NArgs {
  UpQuadArgs<uint8_t, N> l0;  // level 0: labels, g1 = level 1's Gaussians, z = level-0 regions (r1 not read)
  const float* wn1;           // level 1's planar normalised weights, wn1_plane floats per image
  size_t wn1_plane;
  ImgRect z1[N];              // level 1's valid regions
  const float* g2[N];         // level 2's Gaussians
  const float* r2;            // level 2's reconstruction
  int w2, h2;
};

// LDS of up_l10n_kernel, in float4: R1 and each slot's G1 on the level-1
// footprint, then R2 and each slot's G2 on the level-2 footprint.
template <int N>
__host__ __device__ constexpr int up_l10n_lds_float4() {
  return ((N < kQuadSlots ? N : kQuadSlots) + 1) * (kQuadTaps + kL2Taps);
}

// The image set (bit i: image i) of the level-0 tile whose first quad is
// (tqx, tqy): images whose level-0 region z0 meets the tile's 128 x 8 pixels
// or whose level-1 region z1 meets its level-1 footprint.
__host__ __device__ inline unsigned l10n_tile_set(const ImgRect* z0, const ImgRect* z1, int n, int tqx, int tqy) {
  unsigned set = 0;
  for (int i = 0; i < n; ++i) {
    const ImgRect& a = z0[i];
    const ImgRect& c = z1[i];
    const bool m0 = a.x0 <= 2 * tqx + 127 && a.x1 >= 2 * tqx && a.y0 <= 2 * tqy + 7 && a.y1 >= 2 * tqy;
    const bool m1 = c.x0 <= tqx + kQuadTapW - 1 && c.x1 >= tqx && c.y0 <= tqy + kQuadTapH - 1 && c.y1 >= tqy;
    if (m0 || m1) set |= 1u << i;
  }
  return set;
}

template <int C, int N, int S>
__device__ __forceinline__ void up_l10n_slots(const UpL10NArgs<N>& A, const int* idx_rt, int tqx, int tqy, int b,
                                              float4* lds) {
  const UpQuadArgs<uint8_t, N>& a = A.l0;
  constexpr int SA = S > 0 ? S : 1;  // array extent
  int idx[SA];
#pragma unroll
  for (int s = 0; s < S; ++s) idx[s] = idx_rt[s];
  const int lx = (int)threadIdx.x, ly = (int)threadIdx.y;
  const QuadPix q = quad_pixels(a, tqx + lx, tqy + ly, b);
  const int tid = ly * 64 + lx;
  constexpr int kRounds = (kQuadTaps + kWgThreads - 1) / kWgThreads;
  float4* lr1 = lds;                         // R1 on the level-1 footprint
  float4* lg1 = lds + kQuadTaps;             // G1 of slot s at lg1 + s * kQuadTaps
  float4* l2 = lds + (S + 1) * kQuadTaps;    // level 2: R2, then G2 of each slot
  const int bx2 = tqx >> 1, by2 = tqy >> 1;
  // Every load of the workgroup first: the labels (the labelled slot's
  // level-0 pixel depends on them), level 2's footprint, level 1's footprint
  // (each slot's Gaussian and weight), the labelled level-0 pixels.
  int lab[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) lab[e] = static_cast<const uint8_t*>(a.mask0)[q.mpx[e]];
  float4 s2[S + 1];
  {
    const int p = min(tid, kL2Taps - 1);
    const int px = min(bx2 + p % kL2TapW, A.w2 - 1), py = min(by2 + p / kL2TapW, A.h2 - 1);
    const size_t tap = ((size_t)b * A.h2 + py) * A.w2 + px;
    s2[0] = ld_px<C>(A.r2, tap);
#pragma unroll
    for (int s = 0; s < S; ++s) s2[s + 1] = ld_px<C>(A.g2[idx[s]], tap);
  }
  float4 s1[kRounds][SA];
  float w1[kRounds][SA];
  int x1r[kRounds], y1r[kRounds];
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = min(tid + r * kWgThreads, kQuadTaps - 1);
    x1r[r] = min(tqx + p % kQuadTapW, a.w1 - 1);
    y1r[r] = min(tqy + p / kQuadTapW, a.h1 - 1);
    const size_t mp = (size_t)y1r[r] * a.w1 + x1r[r];
    const size_t tap = (size_t)b * a.h1 * a.w1 + mp;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      s1[r][s] = ld_px<C>(a.g1[idx[s]], tap);
      w1[r][s] = A.wn1[(size_t)idx[s] * A.wn1_plane + mp];
    }
  }
  int sl[4];
  uint32_t rv[4];
  label_slot_pixels<C, N, S>(a, idx, q, lab, sl, rv);
  if (tid < kL2Taps) {
#pragma unroll
    for (int k = 0; k <= S; ++k) l2[k * kL2Taps + tid] = s2[k];
  }
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = tid + r * kWgThreads;
    if (p < kQuadTaps) {
#pragma unroll
      for (int s = 0; s < S; ++s) lg1[s * kQuadTaps + p] = s1[r][s];
    }
  }
  __syncthreads();
  // R1 on the footprint (the level-1 launch's quad arithmetic at each position)
#pragma unroll
  for (int r = 0; r < kRounds; ++r) {
    const int p = tid + r * kWgThreads;
    if (p >= kQuadTaps) continue;
    const int x1 = x1r[r], y1 = y1r[r];
    const int e1 = (x1 & 1) | ((y1 & 1) << 1);
    const int t00 = ((y1 >> 1) - by2) * kL2TapW + ((x1 >> 1) - bx2);
    float4 Lv[SA];
    float w[SA];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      Lv[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (in_rect(A.z1[idx[s]], x1, y1)) {
        const float4* gp = l2 + (s + 1) * kL2Taps + t00;
        const float4 gt[4] = {gp[0], gp[1], gp[kL2TapW], gp[kL2TapW + 1]};
        const float4 up = up_quad_lap<C>(e1, gt);
        Lv[s] = make_float4(s1[r][s].x - up.x, s1[r][s].y - up.y, s1[r][s].z - up.z, s1[r][s].w);
      }
      w[s] = w1[r][s];
    }
    const float4 B = blendN_weighted<C, S>(Lv, w);
    const float4* rp = l2 + t00;
    const float4 rt[4] = {rp[0], rp[1], rp[kL2TapW], rp[kL2TapW + 1]};
    const float4 up = up_quad_rec<C>(e1, rt);
    lr1[p] = make_float4(up.x + B.x, up.y + B.y, up.z + B.z, B.w);
  }
  __syncthreads();
  label_level0<C, N, true, S>(a, idx, q, sl, rv, lds, b);
}

// Runs on 64 x 4 threads; lds: up_l10n_lds_float4<N>() float4.
template <int C, int N>
__device__ __forceinline__ void up_l10n_body(const UpL10NArgs<N>& A, Blk blk, float4* lds) {
  const UpQuadArgs<uint8_t, N>& a = A.l0;
  const int t = blk.x;
  const int tqx = a.qx0 + (t % a.tiles_x) * 64, tqy = a.qy0 + (t / a.tiles_x) * 4;
  const unsigned set0 = l10n_tile_set(a.z, A.z1, N, tqx, tqy);  // uniform
  unsigned set = set0;
  int idx[kQuadSlots];
#pragma unroll
  for (int s = 0; s < kQuadSlots; ++s) {
    idx[s] = set ? __builtin_ctz(set) : 0;
    set &= set - 1u;
  }
  switch (__builtin_popcount(set0)) {  // uniform; at most kQuadSlots (the host checks every tile)
    case 0: up_l10n_slots<C, N, 0>(A, idx, tqx, tqy, blk.z, lds); return;
    case 1: up_l10n_slots<C, N, 1>(A, idx, tqx, tqy, blk.z, lds); return;
    case 2: up_l10n_slots<C, N, 2>(A, idx, tqx, tqy, blk.z, lds); return;
    case 3: up_l10n_slots<C, N, 3>(A, idx, tqx, tqy, blk.z, lds); return;
    default:
      if constexpr (N >= 4) up_l10n_slots<C, N, 4>(A, idx, tqx, tqy, blk.z, lds);
      return;
  }
}

template <int C, int N>
__global__ void __launch_bounds__(256) up_l10n_kernel(UpL10NArgs<N> a) {
  __shared__ float4 lds[up_l10n_lds_float4<N>()];
  (void)a;
  up_l10n_body<C, N>(kernel_arg<UpL10NArgs<N>>(), hw_blk(), lds);
}

}  // namespace hmdev
