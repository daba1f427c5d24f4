/*
 * hm_oracle.c -- CPU restatement of hm-cupano's single-frame stitch path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see hm_oracle.h.
 * Built with -ffp-contract=off; every floating-point expression keeps the
 * operand order of the cited reference line so that results are bit-exact
 * with a non-contracting build of the same arithmetic.
 */
#include "hm_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Every per-pixel stage writes each output pixel from its own inputs only, so
 * its rows (and batch entries) run on several threads with identical results;
 * hmo_set_threads() picks how many (1 = the sequential restatement). */
#ifdef _OPENMP
#define HMO_PARALLEL_ROWS _Pragma("omp parallel for collapse(2) schedule(static)")
#define HMO_PARALLEL_FOR _Pragma("omp parallel for schedule(static)")
#else
#define HMO_PARALLEL_ROWS
#define HMO_PARALLEL_FOR
#endif

#define MAXF(a, b) ((a) < (b) ? (b) : (a))
#define IMIN(a, b) ((a) < (b) ? (a) : (b))
#define IMAX(a, b) ((a) > (b) ? (a) : (b))

int hmo_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* src/cuda/cudaUtils.cuh:18-28: clamp to [0,255], then truncate (the +0.5 is
 * commented out in the reference). */
uint8_t hmo_round_to_uchar(float x) {
  if (x <= 0.0f) return 0;
  if (x >= 255.0f) return 255;
  return (uint8_t)x;
}

/* src/cuda/cudaBlend.cu:860-866 (no clamp) and src/cuda/cudaBlendN.cu:675-691
 * (stop when the previous level has a side < 2). */
int hmo_pyramid_sizes(int w, int h, int levels, int clamp_small_sides, int* widths, int* heights) {
  if (levels < 1 || levels > HMO_MAX_LEVELS) return -1;
  widths[0] = w;
  heights[0] = h;
  for (int i = 1; i < levels; ++i) {
    if (clamp_small_sides && (widths[i - 1] < 2 || heights[i - 1] < 2)) return i;
    widths[i] = (widths[i - 1] + 1) / 2;
    heights[i] = (heights[i - 1] + 1) / 2;
  }
  return levels;
}

/* ------------------------------------------------------------------------- */
/* pixel access helpers                                                      */
/* ------------------------------------------------------------------------- */
static size_t pix_index(int b, int x, int y, int w, int h, int c) {
  return (((size_t)b * (size_t)h + (size_t)y) * (size_t)w + (size_t)x) * (size_t)c;
}

/* IEEE binary16 <-> binary32.  Half to float is exact; float to half rounds
 * to nearest, ties to even, like static_cast<__half>(float) in the
 * reference's perform_cast (cudaUtils.cuh:43-47). */
float hmo_f16_to_f32(uint16_t h) {
  const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  const int exp = (h >> 10) & 0x1f;
  uint32_t man = h & 0x3ffu, bits;
  if (exp == 0x1f) {
    bits = sign | 0x7f800000u | (man << 13);
  } else if (exp != 0) {
    bits = sign | ((uint32_t)(exp - 15 + 127) << 23) | (man << 13);
  } else if (man == 0) {
    bits = sign;
  } else { /* subnormal: normalise */
    int e = -14;
    while (!(man & 0x400u)) {
      man <<= 1;
      --e;
    }
    bits = sign | ((uint32_t)(e + 127) << 23) | ((man & 0x3ffu) << 13);
  }
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

uint16_t hmo_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) /* inf, nan (quiet) */
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u | ((ax >> 13) & 0x3ffu) : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to >= 65520: inf */
  if (ax < 0x38800000u) { /* below the smallest normal half: subnormal or zero */
    const int e = (int)(ax >> 23);
    if (e < 102) return sign; /* below half the smallest subnormal (2^-25) */
    const uint32_t m = (ax & 0x7fffffu) | 0x800000u;
    const int shift = 126 - e; /* value = m * 2^(e-150); in units of 2^-24: m >> (126 - e) */
    uint32_t r = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (r & 1u))) ++r;
    return (uint16_t)(sign | r);
  }
  /* normal: keep 10 mantissa bits, round to nearest even (a carry may bump the exponent) */
  uint32_t r = ((ax >> 13) - ((127 - 15) << 10));
  const uint32_t rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (r & 1u))) ++r;
  return (uint16_t)(sign | r);
}

static size_t et_size(int et) { return et == HMO_F32 ? 4 : (et == HMO_F16 ? 2 : 1); }

/* Pixel idx (in elements) of a c-channel image as float; 3 channels read
 * alpha 255 (the 3 -> 4 cast), 1 channel only v[0]. */
static void load_any(const void* base, int et, size_t idx, int c, float v[4]) {
  v[1] = v[2] = 0.0f;
  v[3] = 255.0f;
  for (int k = 0; k < c; ++k) {
    if (et == HMO_F32)
      v[k] = ((const float*)base)[idx + k];
    else if (et == HMO_F16)
      v[k] = hmo_f16_to_f32(((const uint16_t*)base)[idx + k]);
    else
      v[k] = (float)((const uint8_t*)base)[idx + k];
  }
}

/* perform_cast into a dst_c-channel pixel of type et (cudaUtils.cuh:80-213):
 * to uint8 a float or half source is clamped and truncated (round_to_uchar),
 * a uint8 source copied; to half rounds to nearest even; to float is exact;
 * 3 -> 4 gives alpha 255 (v[3]), 4 -> 3 drops alpha. */
static void store_any(void* base, int et, size_t idx, int dst_c, const float v[4], int src_et) {
  for (int k = 0; k < dst_c; ++k) {
    if (et == HMO_F32)
      ((float*)base)[idx + k] = v[k];
    else if (et == HMO_F16)
      ((uint16_t*)base)[idx + k] = hmo_f32_to_f16(v[k]);
    else
      ((uint8_t*)base)[idx + k] = src_et == HMO_U8 ? (uint8_t)v[k] : hmo_round_to_uchar(v[k]);
  }
}

/* ------------------------------------------------------------------------- */
/* remap family                                                              */
/* ------------------------------------------------------------------------- */

/* BatchedRemapKernelExOffset (src/cuda/cudaRemap.cu:61-123) and its ROI form
 * (:129-198) with the launcher's ROI validation (:364-369). The default pixel
 * is T_in{} (all zero), as every pano class passes (cudaPano.inl:176). */
int hmo_remap_offset_t(const void* src, int src_et, int src_w, int src_h, int src_c, void* dst, int dst_et,
                       int dst_w, int dst_h, int dst_c, const uint16_t* map_x, const uint16_t* map_y, int remap_w,
                       int remap_h, int off_x, int off_y, int roi_x, int roi_y, int roi_w, int roi_h,
                       int no_unmapped_write, int batch) {
  if (roi_w < 0) {
    roi_x = 0;
    roi_y = 0;
    roi_w = remap_w;
    roi_h = remap_h;
  } else {
    if (roi_w <= 0 || roi_h <= 0) return 0;
    if (roi_x < 0 || roi_y < 0) return -1;
    if (roi_x + roi_w > remap_w || roi_y + roi_h > remap_h) return -1;
  }
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b) {
    for (int y = 0; y < roi_h; ++y) {
      for (int x = 0; x < roi_w; ++x) {
        const int mx = roi_x + x, my = roi_y + y;
        if (mx < 0 || mx >= remap_w || my < 0 || my >= remap_h) continue;
        const int dx = off_x + mx, dy = off_y + my;
        if (dx < 0 || dx >= dst_w || dy < 0 || dy >= dst_h) continue;
        const size_t di = pix_index(b, dx, dy, dst_w, dst_h, dst_c);
        const size_t mi = (size_t)my * (size_t)remap_w + (size_t)mx;
        const int sx = map_x[mi], sy = map_y[mi];
        if (dst_c == 4 && no_unmapped_write && sx == HMO_UNMAPPED) continue;
        float v[4];
        if (sx < src_w && sy < src_h) {
          load_any(src, src_et, pix_index(b, sx, sy, src_w, src_h, src_c), src_c, v);
          store_any(dst, dst_et, di, dst_c, v, src_et);
        } else if (!no_unmapped_write || sx != HMO_UNMAPPED) {
          v[0] = v[1] = v[2] = 0.0f;
          v[3] = (sx == HMO_UNMAPPED) ? 0.0f : 255.0f; /* alpha of a 4-channel default pixel */
          store_any(dst, dst_et, di, dst_c, v, HMO_U8);
        }
      }
    }
  }
  return 0;
}

int hmo_remap_offset(const uint8_t* src, int src_w, int src_h, int src_c, void* dst, int dst_is_float, int dst_w,
                     int dst_h, int dst_c, const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h,
                     int off_x, int off_y, int roi_x, int roi_y, int roi_w, int roi_h, int no_unmapped_write,
                     int batch) {
  return hmo_remap_offset_t(src, HMO_U8, src_w, src_h, src_c, dst, dst_is_float ? HMO_F32 : HMO_U8, dst_w, dst_h,
                            dst_c, map_x, map_y, remap_w, remap_h, off_x, off_y, roi_x, roi_y, roi_w, roi_h,
                            no_unmapped_write, batch);
}

/* BatchedRemapKernelExOffsetWithDestMap (src/cuda/cudaRemap.cu:204-246):
 * writes only where dest_map[destY*dst_w+destX] == image_index; unmapped or
 * out-of-range sources write the default pixel (all zero, alpha 0 too). */
int hmo_remap_dest_map_t(const void* src, int et, int src_w, int src_h, int c, void* dst, int dst_w, int dst_h,
                         const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h, int off_x, int off_y,
                         int image_index, const uint8_t* dest_map, int batch) {
  const size_t es = et_size(et);
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b) {
    for (int y = 0; y < remap_h; ++y) {
      for (int x = 0; x < remap_w; ++x) {
        const int dx = off_x + x, dy = off_y + y;
        if (dx < 0 || dx >= dst_w || dy < 0 || dy >= dst_h) continue;
        if (dest_map[(size_t)dy * dst_w + dx] != image_index) continue;
        const size_t mi = (size_t)y * remap_w + x;
        const int sx = map_x[mi], sy = map_y[mi];
        uint8_t* d = (uint8_t*)dst + pix_index(b, dx, dy, dst_w, dst_h, c) * es;
        if (sx < src_w && sy < src_h) {
          memcpy(d, (const uint8_t*)src + pix_index(b, sx, sy, src_w, src_h, c) * es, (size_t)c * es);
        } else {
          memset(d, 0, (size_t)c * es);
        }
      }
    }
  }
  return 0;
}

int hmo_remap_dest_map(const uint8_t* src, int src_w, int src_h, int c, uint8_t* dst, int dst_w, int dst_h,
                       const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h, int off_x, int off_y,
                       int image_index, const uint8_t* dest_map, int batch) {
  return hmo_remap_dest_map_t(src, HMO_U8, src_w, src_h, c, dst, dst_w, dst_h, map_x, map_y, remap_w, remap_h, off_x,
                              off_y, image_index, dest_map, batch);
}

/* BatchedRemapHardSeamKernelN (src/cuda/cudaRemap.cu:252-294): per canvas
 * pixel pick the seam's image, gather; sentinel / out-of-range leave the
 * destination untouched. */
int hmo_remap_hard_seam_n_t(int n, const void* const* srcs, int et, const int* src_w, const int* src_h, int c,
                            const uint16_t* const* map_x, const uint16_t* const* map_y, const int* off_x,
                            const int* off_y, const int* map_w, const int* map_h, const uint8_t* dest_map, void* dst,
                            int dst_w, int dst_h, int batch) {
  const size_t es = et_size(et);
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b) {
    for (int y = 0; y < dst_h; ++y) {
      for (int x = 0; x < dst_w; ++x) {
        const int i = dest_map[(size_t)y * dst_w + x];
        if (i < 0 || i >= n) continue;
        const int rx = x - off_x[i], ry = y - off_y[i];
        if (rx < 0 || ry < 0 || rx >= map_w[i] || ry >= map_h[i]) continue;
        const size_t mi = (size_t)ry * map_w[i] + rx;
        const int sx = map_x[i][mi], sy = map_y[i][mi];
        if (sx == HMO_UNMAPPED) continue;
        if (sx < src_w[i] && sy < src_h[i]) {
          memcpy((uint8_t*)dst + pix_index(b, x, y, dst_w, dst_h, c) * es,
                 (const uint8_t*)srcs[i] + pix_index(b, sx, sy, src_w[i], src_h[i], c) * es, (size_t)c * es);
        }
      }
    }
  }
  return 0;
}

int hmo_remap_hard_seam_n(int n, const uint8_t* const* srcs, const int* src_w, const int* src_h, int c,
                          const uint16_t* const* map_x, const uint16_t* const* map_y, const int* off_x,
                          const int* off_y, const int* map_w, const int* map_h, const uint8_t* dest_map, uint8_t* dst,
                          int dst_w, int dst_h, int batch) {
  return hmo_remap_hard_seam_n_t(n, (const void* const*)srcs, HMO_U8, src_w, src_h, c, map_x, map_y, off_x, off_y,
                                 map_w, map_h, dest_map, dst, dst_w, dst_h, batch);
}

/* copyRoiKernelBatched (src/cuda/cudaMakeFull.cu:40-69) with perform_cast:
 * u8->float exact, float->u8 truncating (cudaUtils.cuh:118-127), float->float
 * and u8->u8 copies, half as float; channel conversion as in
 * cudaUtils.cuh:141-183. */
int hmo_copy_roi_t(const void* src, int src_et, int src_w, int src_h, int src_c, int region_w, int region_h,
                   int src_roi_x, int src_roi_y, void* dst, int dst_et, int dst_w, int dst_h, int dst_c, int off_x,
                   int off_y, int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b) {
    for (int y = 0; y < region_h; ++y) {
      for (int x = 0; x < region_w; ++x) {
        const int sx = src_roi_x + x, sy = src_roi_y + y;
        if (!(sx < src_w && sy < src_h)) continue;
        const int dx = off_x + x, dy = off_y + y;
        if (!(dx < dst_w && dy < dst_h)) continue;
        float v[4];
        load_any(src, src_et, pix_index(b, sx, sy, src_w, src_h, src_c), src_c, v);
        store_any(dst, dst_et, pix_index(b, dx, dy, dst_w, dst_h, dst_c), dst_c, v, src_et);
      }
    }
  }
  return 0;
}

int hmo_copy_roi(const void* src, int src_is_float, int src_w, int src_h, int src_c, int region_w, int region_h,
                 int src_roi_x, int src_roi_y, void* dst, int dst_is_float, int dst_w, int dst_h, int dst_c,
                 int off_x, int off_y, int batch) {
  return hmo_copy_roi_t(src, src_is_float ? HMO_F32 : HMO_U8, src_w, src_h, src_c, region_w, region_h, src_roi_x,
                        src_roi_y, dst, dst_is_float ? HMO_F32 : HMO_U8, dst_w, dst_h, dst_c, off_x, off_y, batch);
}

/* ------------------------------------------------------------------------- */
/* pyramid stages shared by the 2-, 3- and N-image blends                    */
/* ------------------------------------------------------------------------- */

/* FusedBatchedDownsampleKernel, per image (src/cuda/cudaBlend.cu:46-144;
 * same arithmetic in cudaBlendN.cu:53-113 and cudaBlend3.cu:100-215):
 * RGB mean of the in-range 2x2 taps whose alpha != 0 (dy outer, dx inner),
 * 0 if none; alpha = max over all in-range taps. */
void hmo_downsample(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int c, int batch) {
  const int sum_c = c < 3 ? c : 3;
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < out_h; ++y)
      for (int x = 0; x < out_w; ++x) {
        float sums[3] = {0.0f, 0.0f, 0.0f};
        int count = 0;
        float alpha = 0.0f;
        for (int dy = 0; dy < 2; ++dy)
          for (int dx = 0; dx < 2; ++dx) {
            const int ix = 2 * x + dx, iy = 2 * y + dy;
            if (ix >= in_w || iy >= in_h) continue;
            const float* p = in + pix_index(b, ix, iy, in_w, in_h, c);
            int keep = 1;
            if (c == 4) {
              alpha = MAXF(alpha, p[3]);
              keep = (p[3] != 0.0f);
            }
            if (keep) {
              for (int k = 0; k < sum_c; ++k) sums[k] += p[k];
              ++count;
            }
          }
        float* o = out + pix_index(b, x, y, out_w, out_h, c);
        for (int k = 0; k < sum_c; ++k) o[k] = count > 0 ? sums[k] / (float)count : 0.0f;
        if (c == 4) o[3] = alpha;
      }
}

/* BatchedDownsampleKernelMask (src/cuda/cudaBlend.cu:268-296), generalised
 * to M interleaved channels as FusedBatchedDownsampleMaskN
 * (cudaBlendN.cu:19-48) and FusedBatchedDownsampleMask3 (cudaBlend3.cu:220-262). */
void hmo_downsample_mask(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int m) {
  HMO_PARALLEL_FOR
  for (int y = 0; y < out_h; ++y)
    for (int x = 0; x < out_w; ++x) {
      float sums[HMO_MAX_IMAGES];
      int count = 0;
      for (int k = 0; k < m; ++k) sums[k] = 0.0f;
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
}

/* Top-left aligned upsample sample positions shared by the Laplacian and
 * reconstruct kernels (cudaBlend.cu:383-390 and :613-623). */
typedef struct {
  int gxi, gyi, gxi1, gyi1;
  float dx, dy;
} up_taps;

static up_taps make_taps(int x, int y, int low_w, int low_h, int clamp_base) {
  up_taps t;
  const float gx = (float)x / 2.0f, gy = (float)y / 2.0f;
  int gxi = (int)gx, gyi = (int)gy; /* floorf of a non-negative value */
  if (clamp_base) {
    gxi = IMAX(0, IMIN(gxi, low_w - 1));
    gyi = IMAX(0, IMIN(gyi, low_h - 1));
  }
  t.gxi = gxi;
  t.gyi = gyi;
  t.gxi1 = IMIN(gxi + 1, low_w - 1);
  t.gyi1 = IMIN(gyi + 1, low_h - 1);
  t.dx = gx - (float)gxi;
  t.dy = gy - (float)gyi;
  return t;
}

/* Alpha-aware bilinear value of channel k (cudaBlend.cu:416-444 and the
 * try_add lambda of :649-666): taps 00,10,01,11 in that order, taps whose
 * alpha is 0 are skipped (4-channel only), value = sumV / sumW or 0. */
static float up_alpha_aware(const float* low, int low_w, int low_h, int c, int b, const up_taps* t, int k) {
  const float w00 = (1.0f - t->dx) * (1.0f - t->dy);
  const float w10 = t->dx * (1.0f - t->dy);
  const float w01 = (1.0f - t->dx) * t->dy;
  const float w11 = t->dx * t->dy;
  const float* p00 = low + pix_index(b, t->gxi, t->gyi, low_w, low_h, c);
  const float* p10 = low + pix_index(b, t->gxi1, t->gyi, low_w, low_h, c);
  const float* p01 = low + pix_index(b, t->gxi, t->gyi1, low_w, low_h, c);
  const float* p11 = low + pix_index(b, t->gxi1, t->gyi1, low_w, low_h, c);
  float sum_w = 0.0f, sum_v = 0.0f;
  if (c != 4 || p00[3] != 0.0f) { sum_w += w00; sum_v += p00[k] * w00; }
  if (c != 4 || p10[3] != 0.0f) { sum_w += w10; sum_v += p10[k] * w10; }
  if (c != 4 || p01[3] != 0.0f) { sum_w += w01; sum_v += p01[k] * w01; }
  if (c != 4 || p11[3] != 0.0f) { sum_w += w11; sum_v += p11[k] * w11; }
  return (sum_w > 0.0f) ? (sum_v / sum_w) : 0.0f;
}

/* BatchedComputeLaplacianKernel (src/cuda/cudaBlend.cu:356-456): 4-channel
 * uses the alpha-aware upsample, 3-channel the plain two-stage lerp
 * (:446-449); alpha is copied from the high level. */
void hmo_laplacian(const float* high, int hw, int hh, const float* low, int lw, int lh, float* lap, int c,
                            int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < hh; ++y)
      for (int x = 0; x < hw; ++x) {
        const up_taps t = make_taps(x, y, lw, lh, 0);
        const float* hp = high + pix_index(b, x, y, hw, hh, c);
        float* o = lap + pix_index(b, x, y, hw, hh, c);
        for (int k = 0; k < c; ++k) {
          if (c == 4 && k == 3) {
            o[3] = hp[3];
            continue;
          }
          float up;
          if (c == 4) {
            up = up_alpha_aware(low, lw, lh, c, b, &t, k);
          } else {
            const float v00 = low[pix_index(b, t.gxi, t.gyi, lw, lh, c) + k];
            const float v10 = low[pix_index(b, t.gxi1, t.gyi, lw, lh, c) + k];
            const float v01 = low[pix_index(b, t.gxi, t.gyi1, lw, lh, c) + k];
            const float v11 = low[pix_index(b, t.gxi1, t.gyi1, lw, lh, c) + k];
            const float v0 = v00 * (1.0f - t.dx) + v10 * t.dx;
            const float v1 = v01 * (1.0f - t.dx) + v11 * t.dx;
            up = v0 * (1.0f - t.dy) + v1 * t.dy;
          }
          o[k] = hp[k] - up;
        }
      }
}

/* BatchedReconstructKernel (src/cuda/cudaBlend.cu:581-672): clamped base
 * taps, alpha-aware upsample of the lower level plus the blended Laplacian;
 * alpha copied from the Laplacian. */
void hmo_reconstruct(const float* low, int lw, int lh, const float* lap, int hw, int hh, float* out, int c,
                              int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < hh; ++y)
      for (int x = 0; x < hw; ++x) {
        const up_taps t = make_taps(x, y, lw, lh, 1);
        const float* lp = lap + pix_index(b, x, y, hw, hh, c);
        float* o = out + pix_index(b, x, y, hw, hh, c);
        for (int k = 0; k < c; ++k) {
          if (c == 4 && k == 3) {
            o[3] = lp[3];
            continue;
          }
          const float up = up_alpha_aware(low, lw, lh, c, b, &t, k);
          o[k] = up + lp[k];
        }
      }
}

/* BatchedBlendKernel (src/cuda/cudaBlend.cu:520-576). */
static void blend2_level(const float* l1, const float* l2, const float* mask, float* out, int w, int h, int c,
                         int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const size_t i = pix_index(b, x, y, w, h, c);
        const float m = mask[(size_t)y * w + x];
        const float mm1 = 1.0f - m;
        const int max_c = c < 3 ? c : 3;
        if (c == 4) {
          if (l1[i + 3] == 0.0f) {
            for (int k = 0; k < 4; ++k) out[i + k] = l2[i + k];
          } else if (l2[i + 3] == 0.0f) {
            for (int k = 0; k < 4; ++k) out[i + k] = l1[i + k];
          } else {
            for (int k = 0; k < max_c; ++k) out[i + k] = m * l1[i + k] + mm1 * l2[i + k];
            out[i + 3] = l2[i + 3];
          }
        } else {
          for (int k = 0; k < max_c; ++k) out[i + k] = m * l1[i + k] + mm1 * l2[i + k];
        }
      }
}

/* BatchedBlendKernelN (src/cuda/cudaBlendN.cu:199-288). */
static void blend_n_level(int n, const float* const* laps, const float* mask, float* out, int w, int h, int c,
                          int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const size_t pix = (size_t)y * w + x;
        const size_t base = pix_index(b, x, y, w, h, c);
        float wt[HMO_MAX_IMAGES];
        float sum_w = 0.0f;
        for (int i = 0; i < n; ++i) {
          wt[i] = mask[pix * n + i];
          sum_w += wt[i];
        }
        if (sum_w > 0.0f)
          for (int i = 0; i < n; ++i) wt[i] /= sum_w;
        if (c == 4) {
          float sum_w2 = 0.0f, max_alpha = 0.0f;
          int max_idx = -1;
          for (int i = 0; i < n; ++i) {
            const float a = laps[i][base + 3];
            if (a > max_alpha) {
              max_alpha = a;
              max_idx = i;
            }
            if (a == 0.0f) wt[i] = 0.0f;
            sum_w2 += wt[i];
          }
          if (sum_w2 > 0.0f) {
            for (int i = 0; i < n; ++i) wt[i] /= sum_w2;
          } else {
            if (max_idx >= 0 && max_alpha != 0.0f) {
              for (int k = 0; k < 4; ++k) out[base + k] = laps[max_idx][base + k];
            } else {
              for (int k = 0; k < 4; ++k) out[base + k] = 0.0f;
            }
            continue;
          }
        }
        for (int k = 0; k < c; ++k) {
          if (c == 4 && k == 3) {
            float a_out = 0.0f;
            for (int i = 0; i < n; ++i) {
              if (wt[i] == 0.0f) continue;
              a_out = MAXF(a_out, laps[i][base + 3]);
            }
            out[base + 3] = a_out;
          } else {
            float acc = 0.0f;
            for (int i = 0; i < n; ++i) acc += wt[i] * laps[i][base + k];
            out[base + k] = acc;
          }
        }
      }
}

/* BatchedBlendKernel3 (src/cuda/cudaBlend3.cu:376-531). */
static void blend3_level(const float* l1, const float* l2, const float* l3, const float* mask, float* out, int w,
                         int h, int c, int batch) {
  HMO_PARALLEL_ROWS
  for (int b = 0; b < batch; ++b)
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const float* mp = mask + ((size_t)y * w + x) * 3;
        float m1 = mp[0], m2 = mp[1], m3 = mp[2];
        const size_t idx = pix_index(b, x, y, w, h, c);
        if (c == 4) {
          const float a1 = l1[idx + 3], a2 = l2[idx + 3], a3 = l3[idx + 3];
          if (a1 == 0.0f) m1 = 0.0f;
          if (a2 == 0.0f) m2 = 0.0f;
          if (a3 == 0.0f) m3 = 0.0f;
          const float sum = m1 + m2 + m3;
          if (sum > 0.0f) {
            m1 /= sum;
            m2 /= sum;
            m3 /= sum;
          } else {
            const float* src = NULL;
            float amax = 0.0f;
            if (a1 > amax) { amax = a1; src = l1 + idx; }
            if (a2 > amax) { amax = a2; src = l2 + idx; }
            if (a3 > amax) { amax = a3; src = l3 + idx; }
            for (int k = 0; k < c; ++k) out[idx + k] = src ? src[k] : 0.0f;
            continue;
          }
        }
        const int bc = c < 3 ? c : 3;
        for (int k = 0; k < bc; ++k) {
          const float v1 = l1[idx + k], v2 = l2[idx + k], v3 = l3[idx + k];
          out[idx + k] = m1 * v1 + m2 * v2 + m3 * v3;
        }
        if (c == 4) {
          float a_out = 0.0f;
          if (m1 > 0.0f) a_out = MAXF(a_out, l1[idx + 3]);
          if (m2 > 0.0f) a_out = MAXF(a_out, l2[idx + 3]);
          if (m3 > 0.0f) a_out = MAXF(a_out, l3[idx + 3]);
          out[idx + 3] = a_out;
        }
      }
}

/* ------------------------------------------------------------------------- */
/* generic pyramid blend driver (kind: 2 = 2-image, 3 = 3-image, 0 = N)      */
/* ------------------------------------------------------------------------- */
static void dump_copy(float* dst, const float* src, size_t n) {
  if (dst && src) memcpy(dst, src, n * sizeof(float));
}

static int blend_driver(int kind, int n, const float* const* imgs, const float* mask, float* out, int w, int h,
                        int c, int levels_in, int batch, hmo_dump* dump) {
  if (n < 2 || n > HMO_MAX_IMAGES || (c != 3 && c != 4) || w < 1 || h < 1 || batch < 1) return -1;
  int widths[HMO_MAX_LEVELS], heights[HMO_MAX_LEVELS];
  const int levels = hmo_pyramid_sizes(w, h, levels_in, kind != 2, widths, heights);
  if (levels < 1) return -1;
  const int m = (kind == 2) ? 1 : n;

  float* gauss[HMO_MAX_IMAGES][HMO_MAX_LEVELS];
  float* lap[HMO_MAX_IMAGES][HMO_MAX_LEVELS];
  float* maskp[HMO_MAX_LEVELS];
  float* blend[HMO_MAX_LEVELS];
  float* recon[HMO_MAX_LEVELS];
  size_t npx[HMO_MAX_LEVELS];
  for (int l = 0; l < levels; ++l) {
    npx[l] = (size_t)widths[l] * heights[l];
    for (int i = 0; i < n; ++i) {
      gauss[i][l] = l ? (float*)malloc(npx[l] * c * batch * sizeof(float)) : (float*)imgs[i];
      lap[i][l] = (float*)malloc(npx[l] * c * batch * sizeof(float));
    }
    maskp[l] = l ? (float*)malloc(npx[l] * m * sizeof(float)) : (float*)mask;
    blend[l] = (float*)malloc(npx[l] * c * batch * sizeof(float));
    recon[l] = l ? (float*)malloc(npx[l] * c * batch * sizeof(float)) : out;
  }

  /* 1. Gaussian pyramids and mask pyramid (cudaBlend.cu:894-924). */
  for (int l = 0; l + 1 < levels; ++l) {
    for (int i = 0; i < n; ++i)
      hmo_downsample(gauss[i][l], widths[l], heights[l], gauss[i][l + 1], widths[l + 1], heights[l + 1], c, batch);
    hmo_downsample_mask(maskp[l], widths[l], heights[l], maskp[l + 1], widths[l + 1], heights[l + 1], m);
  }
  /* 2. Laplacians (:926-954); coarsest level copies the Gaussian (:955-967). */
  for (int l = 0; l + 1 < levels; ++l)
    for (int i = 0; i < n; ++i)
      hmo_laplacian(gauss[i][l], widths[l], heights[l], gauss[i][l + 1], widths[l + 1], heights[l + 1],
                      lap[i][l], c, batch);
  const int last = levels - 1;
  for (int i = 0; i < n; ++i) memcpy(lap[i][last], gauss[i][last], npx[last] * c * batch * sizeof(float));
  /* 3. Blend every level (:969-985). */
  for (int l = 0; l < levels; ++l) {
    if (kind == 2)
      blend2_level(lap[0][l], lap[1][l], maskp[l], blend[l], widths[l], heights[l], c, batch);
    else if (kind == 3)
      blend3_level(lap[0][l], lap[1][l], lap[2][l], maskp[l], blend[l], widths[l], heights[l], c, batch);
    else {
      const float* lp[HMO_MAX_IMAGES];
      for (int i = 0; i < n; ++i) lp[i] = lap[i][l];
      blend_n_level(n, lp, maskp[l], blend[l], widths[l], heights[l], c, batch);
    }
  }
  /* 4. Reconstruct from the coarsest level (:987-1044). */
  memcpy(recon[last], blend[last], npx[last] * c * batch * sizeof(float));
  for (int l = last - 1; l >= 0; --l)
    hmo_reconstruct(recon[l + 1], widths[l + 1], heights[l + 1], blend[l], widths[l], heights[l], recon[l], c,
                      batch);

  if (dump) {
    dump->levels = levels;
    for (int l = 0; l < levels; ++l) {
      dump->widths[l] = widths[l];
      dump->heights[l] = heights[l];
      for (int i = 0; i < n; ++i) {
        dump_copy(dump->gauss[i][l], gauss[i][l], npx[l] * c * batch);
        dump_copy(dump->lap[i][l], lap[i][l], npx[l] * c * batch);
      }
      dump_copy(dump->mask[l], maskp[l], npx[l] * m);
      dump_copy(dump->blend[l], blend[l], npx[l] * c * batch);
      dump_copy(dump->reconstruct[l], recon[l], npx[l] * c * batch);
    }
  }
  for (int l = 0; l < levels; ++l) {
    for (int i = 0; i < n; ++i) {
      if (l) free(gauss[i][l]);
      free(lap[i][l]);
    }
    if (l) {
      free(maskp[l]);
      free(recon[l]);
    }
    free(blend[l]);
  }
  return 0;
}

int hmo_blend2(const float* img1, const float* img2, const float* mask, float* out, int w, int h, int c, int levels,
               int batch, hmo_dump* dump) {
  const float* imgs[2] = {img1, img2};
  return blend_driver(2, 2, imgs, mask, out, w, h, c, levels, batch, dump);
}

int hmo_blend_n(int n, const float* const* imgs, const float* mask, float* out, int w, int h, int c, int levels,
                int batch, hmo_dump* dump) {
  return blend_driver(0, n, imgs, mask, out, w, h, c, levels, batch, dump);
}

int hmo_blend3(const float* img1, const float* img2, const float* img3, const float* mask, float* out, int w, int h,
               int c, int levels, int batch, hmo_dump* dump) {
  const float* imgs[3] = {img1, img2, img3};
  return blend_driver(3, 3, imgs, mask, out, w, h, c, levels, batch, dump);
}

/* ------------------------------------------------------------------------- */
/* mask padding: CanvasManager::convertMaskMat (canvasManager.cpp:63-101)     */
/* = cv::copyMakeBorder(BORDER_REPLICATE) on the right/bottom                */
/* ------------------------------------------------------------------------- */
int hmo_pad_replicate(const uint8_t* seam, int sw, int sh, uint8_t* out, int cw, int ch) {
  /* canvasManager.cpp:69-70 asserts the seam is no larger than the canvas. */
  if (sw > cw || sh > ch || sw < 1 || sh < 1) return -1;
  HMO_PARALLEL_FOR
  for (int y = 0; y < ch; ++y)
    for (int x = 0; x < cw; ++x) out[(size_t)y * cw + x] = seam[(size_t)IMIN(y, sh - 1) * sw + IMIN(x, sw - 1)];
  return 0;
}

static uint8_t* pad_replicate(const uint8_t* seam, int sw, int sh, int cw, int ch) {
  uint8_t* out = (uint8_t*)malloc((size_t)cw * ch);
  if (hmo_pad_replicate(seam, sw, sh, out, cw, ch) != 0) {
    free(out);
    return NULL;
  }
  return out;
}

/* ------------------------------------------------------------------------- */
/* 2-image pano                                                              */
/* ------------------------------------------------------------------------- */
#define OVERLAP_PAD 128 /* CanvasManager default overlap_pad (canvasManager.h:65) */

/* ControlMasks::canvas_width/height (src/pano/controlMasks.cpp:265-271). */
int hmo_pano2_canvas_size(const hmo_pano2_cfg* cfg, int* canvas_w, int* canvas_h) {
  *canvas_w = IMAX(cfg->pos_x[0] + cfg->map_w[0], cfg->pos_x[1] + cfg->map_w[1]);
  *canvas_h = IMAX(cfg->pos_y[0] + cfg->map_h[0], cfg->pos_y[1] + cfg->map_h[1]);
  return 0;
}

int hmo_pano2_process(const hmo_pano2_cfg* cfg, const void* in1, const void* in2, void* canvas, hmo_dump* dump) {
  const int c = cfg->channels, batch = cfg->batch;
  const int pe = cfg->pipeline_f32 ? HMO_F32 : HMO_U8; /* element type of inputs and canvas */
  if ((c != 3 && c != 4) || batch < 1 || cfg->levels < 0) return -1;
  int cw, ch;
  hmo_pano2_canvas_size(cfg, &cw, &ch);
  const int hard = cfg->levels == 0;
  const int minimize = cfg->minimize_blend && !hard;
  /* CanvasManager::updateMinimizeBlend (canvasManager.cpp:16-61). */
  const int x1 = cfg->pos_x[0], y1 = cfg->pos_y[0], x2 = cfg->pos_x[1], y2 = cfg->pos_y[1];
  const int width_1 = cfg->map_w[0];
  const int overlap = width_1 - x2;
  const int blend_width = overlap + 2 * OVERLAP_PAD;
  int roi1_x = 0, roi_w = 0, rm2_xpos = 0;
  if (minimize) {
    if (overlap <= 0 || x2 - OVERLAP_PAD < 0 || width_1 + OVERLAP_PAD > cw) return -1;
    roi1_x = x2 - OVERLAP_PAD - x1;
    roi_w = blend_width - OVERLAP_PAD;
    rm2_xpos = x1 + OVERLAP_PAD;
  }
  /* convertMaskMat (canvasManager.cpp:63-101). */
  uint8_t* padded = pad_replicate(cfg->seam, cfg->seam_w, cfg->seam_h, cw, ch);
  if (!padded) return -1;
  const size_t cpx = (size_t)cw * ch;
  memset(canvas, 0, cpx * c * batch * et_size(pe)); /* cudaPano.inl:186-187 */

  if (hard) {
    /* cudaPano.inl:262-276 and :345-359 (indices inverted at load time). */
    hmo_remap_dest_map_t(in1, pe, cfg->src_w[0], cfg->src_h[0], c, canvas, cw, ch, cfg->map_x[0], cfg->map_y[0],
                         cfg->map_w[0], cfg->map_h[0], x1, y1, 1, padded, batch);
    hmo_remap_dest_map_t(in2, pe, cfg->src_w[1], cfg->src_h[1], c, canvas, cw, ch, cfg->map_x[1], cfg->map_y[1],
                         cfg->map_w[1], cfg->map_h[1], x2, y2, 0, padded, batch);
    free(padded);
    return 0;
  }

  const int bx0 = minimize ? x2 - OVERLAP_PAD : 0;
  const int bw = minimize ? blend_width : cw;
  const int bh = ch;
  float* mask = (float*)malloc((size_t)bw * bh * sizeof(float));
  HMO_PARALLEL_FOR
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) mask[(size_t)y * bw + x] = (float)padded[(size_t)y * cw + bx0 + x];
  free(padded);
  const size_t fpx = (size_t)bw * bh * c * batch;
  float* full1 = (float*)calloc(fpx, sizeof(float));
  float* full2 = (float*)calloc(fpx, sizeof(float));

  if (minimize) {
    /* cudaPano.inl:210-237 */
    hmo_remap_offset_t(in1, pe, cfg->src_w[0], cfg->src_h[0], c, canvas, pe, cw, ch, c, cfg->map_x[0],
                       cfg->map_y[0], cfg->map_w[0], cfg->map_h[0], x1, y1, 0, 0, -1, -1, 0, batch);
    memset(full1, 0, fpx * sizeof(float));
    hmo_copy_roi_t(canvas, pe, cw, ch, c, roi_w, bh, roi1_x, 0, full1, HMO_F32, bw, bh, c, x1, 0, batch);
    /* cudaPano.inl:291-318 */
    hmo_remap_offset_t(in2, pe, cfg->src_w[1], cfg->src_h[1], c, canvas, pe, cw, ch, c, cfg->map_x[1],
                       cfg->map_y[1], cfg->map_w[1], cfg->map_h[1], x2, y2, 0, 0, -1, -1, 0, batch);
    memset(full2, 0, fpx * sizeof(float));
    hmo_copy_roi_t(canvas, pe, cw, ch, c, roi_w, bh, x2, 0, full2, HMO_F32, bw, bh, c, rm2_xpos, 0, batch);
  } else {
    /* cudaPano.inl:242-255 and :323-336 */
    hmo_remap_offset_t(in1, pe, cfg->src_w[0], cfg->src_h[0], c, full1, HMO_F32, bw, bh, c, cfg->map_x[0],
                       cfg->map_y[0], cfg->map_w[0], cfg->map_h[0], x1, y1, 0, 0, -1, -1, 0, batch);
    hmo_remap_offset_t(in2, pe, cfg->src_w[1], cfg->src_h[1], c, full2, HMO_F32, bw, bh, c, cfg->map_x[1],
                       cfg->map_y[1], cfg->map_w[1], cfg->map_h[1], x2, y2, 0, 0, -1, -1, 0, batch);
  }
  /* cudaPano.inl:384-392 (output aliases full1) and :401-411 (write-back). */
  float* blended = (float*)malloc(fpx * sizeof(float));
  int rc = hmo_blend2(full1, full2, mask, blended, bw, bh, c, cfg->levels, batch, dump);
  if (rc == 0)
    hmo_copy_roi_t(blended, HMO_F32, bw, bh, c, bw, bh, 0, 0, canvas, pe, cw, ch, c, minimize ? x2 - OVERLAP_PAD : 0, 0,
                   batch);
  free(blended);
  free(full1);
  free(full2);
  free(mask);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* N-image pano                                                              */
/* ------------------------------------------------------------------------- */
typedef struct {
  int x, y, w, h;
} rect_t;

/* detailN helpers (src/pano/cudaPanoN.inl:20-105). */
static int pyramid_margin(int levels) { return levels <= 0 ? 0 : 1 << IMIN(levels, 30); }
static int pyramid_alignment(int levels) { return levels <= 1 ? 1 : 1 << IMIN(levels - 1, 30); }

static int seam_boundary_bbox(const uint8_t* s, int w, int h, rect_t* out) {
  if (w <= 1 || h <= 1) return 0;
  int min_x = w, min_y = h, max_x = -1, max_y = -1;
#define UPD(xx, yy)              \
  do {                           \
    min_x = IMIN(min_x, (xx));   \
    min_y = IMIN(min_y, (yy));   \
    max_x = IMAX(max_x, (xx));   \
    max_y = IMAX(max_y, (yy));   \
  } while (0)
  for (int y = 0; y < h; ++y) {
    const uint8_t* row = s + (size_t)y * w;
    const uint8_t* down = (y + 1 < h) ? row + w : NULL;
    for (int x = 0; x < w; ++x) {
      const uint8_t v = row[x];
      if (x + 1 < w && v != row[x + 1]) {
        UPD(x, y);
        UPD(x + 1, y);
      }
      if (down && v != down[x]) {
        UPD(x, y);
        UPD(x, y + 1);
      }
    }
  }
#undef UPD
  if (max_x < 0 || max_y < 0) return 0;
  out->x = min_x;
  out->y = min_y;
  out->w = max_x - min_x + 1;
  out->h = max_y - min_y + 1;
  return 1;
}

static rect_t expand_and_clamp(rect_t r, int pad, int max_w, int max_h) {
  rect_t e = {0, 0, 0, 0};
  if (r.w <= 0 || r.h <= 0) return e;
  const int x0 = IMAX(0, r.x - pad), y0 = IMAX(0, r.y - pad);
  const int x1 = IMIN(max_w, r.x + r.w + pad), y1 = IMIN(max_h, r.y + r.h + pad);
  e.x = x0;
  e.y = y0;
  e.w = IMAX(0, x1 - x0);
  e.h = IMAX(0, y1 - y0);
  return e;
}

static rect_t align_and_clamp(rect_t r, int align, int max_w, int max_h) {
  rect_t e = {0, 0, 0, 0};
  if (r.w <= 0 || r.h <= 0) return e;
  if (align <= 1) return r;
  const int x0 = IMAX(0, (r.x / align) * align), y0 = IMAX(0, (r.y / align) * align);
  const int x1 = IMIN(max_w, ((r.x + r.w + align - 1) / align) * align);
  const int y1 = IMIN(max_h, ((r.y + r.h + align - 1) / align) * align);
  e.x = x0;
  e.y = y0;
  e.w = IMAX(0, x1 - x0);
  e.h = IMAX(0, y1 - y0);
  return e;
}

/* cv::Rect operator& */
static rect_t rect_and(rect_t a, rect_t b) {
  rect_t e = {0, 0, 0, 0};
  const int x0 = IMAX(a.x, b.x), y0 = IMAX(a.y, b.y);
  const int x1 = IMIN(a.x + a.w, b.x + b.w), y1 = IMIN(a.y + a.h, b.y + b.h);
  if (x1 - x0 <= 0 || y1 - y0 <= 0) return e;
  e.x = x0;
  e.y = y0;
  e.w = x1 - x0;
  e.h = y1 - y0;
  return e;
}

/* ControlMasksN::canvas_width/height (src/pano/controlMasksN.cpp:205-219). */
int hmo_panoN_canvas_size(const hmo_panoN_cfg* cfg, int* canvas_w, int* canvas_h) {
  int w = 0, h = 0;
  for (int i = 0; i < cfg->n; ++i) {
    w = IMAX(w, cfg->pos_x[i] + cfg->map_w[i]);
    h = IMAX(h, cfg->pos_y[i] + cfg->map_h[i]);
  }
  *canvas_w = w;
  *canvas_h = h;
  return 0;
}

/* The ROI choice of the CudaStitchPanoN ctor (cudaPanoN.inl:155-190). */
static int panoN_rois(const hmo_panoN_cfg* cfg, const uint8_t* padded, int cw, int ch, rect_t* write_roi,
                      rect_t* blend_roi) {
  rect_t zero = {0, 0, 0, 0};
  *write_roi = zero;
  *blend_roi = zero;
  rect_t bb;
  if (!seam_boundary_bbox(padded, cw, ch, &bb)) return 0;
  *write_roi = expand_and_clamp(bb, OVERLAP_PAD, cw, ch);
  *blend_roi = expand_and_clamp(*write_roi, pyramid_margin(cfg->levels), cw, ch);
  *blend_roi = align_and_clamp(*blend_roi, pyramid_alignment(cfg->levels), cw, ch);
  return 1;
}

int hmo_panoN_rois(const hmo_panoN_cfg* cfg, int* write_roi, int* blend_roi) {
  int cw, ch;
  hmo_panoN_canvas_size(cfg, &cw, &ch);
  uint8_t* padded = pad_replicate(cfg->seam, cfg->seam_w, cfg->seam_h, cw, ch);
  if (!padded) return -1;
  rect_t wr, br;
  panoN_rois(cfg, padded, cw, ch, &wr, &br);
  free(padded);
  write_roi[0] = wr.x;
  write_roi[1] = wr.y;
  write_roi[2] = wr.w;
  write_roi[3] = wr.h;
  blend_roi[0] = br.x;
  blend_roi[1] = br.y;
  blend_roi[2] = br.w;
  blend_roi[3] = br.h;
  return 0;
}

int hmo_panoN_process(const hmo_panoN_cfg* cfg, const void* const* inputs, void* canvas, hmo_dump* dump) {
  const int n = cfg->n, c = cfg->channels, batch = cfg->batch;
  const int pe = cfg->pipeline_f32 ? HMO_F32 : HMO_U8; /* element type of inputs and canvas */
  const size_t ces = et_size(pe);
  if (n < 2 || n > HMO_MAX_IMAGES || (c != 3 && c != 4) || batch < 1 || cfg->levels < 0) return -1;
  int cw, ch;
  hmo_panoN_canvas_size(cfg, &cw, &ch);
  const int minimize = cfg->minimize_blend && cfg->levels > 0;
  uint8_t* padded = pad_replicate(cfg->seam, cfg->seam_w, cfg->seam_h, cw, ch);
  if (!padded) return -1;
  const size_t cpx = (size_t)cw * ch;

  if (cfg->levels == 0) {
    /* cudaPanoN.inl:508-541 */
    memset(canvas, 0, cpx * c * batch * ces);
    hmo_remap_hard_seam_n_t(n, inputs, pe, cfg->src_w, cfg->src_h, c, cfg->map_x, cfg->map_y, cfg->pos_x, cfg->pos_y,
                            cfg->map_w, cfg->map_h, padded, canvas, cw, ch, batch);
    free(padded);
    return 0;
  }

  rect_t write_roi = {0, 0, 0, 0}, blend_roi = {0, 0, 0, 0};
  if (minimize) panoN_rois(cfg, padded, cw, ch, &write_roi, &blend_roi);
  const int use_min = minimize && blend_roi.w > 0 && blend_roi.h > 0 && write_roi.w > 0 && write_roi.h > 0;
  const int bw = (minimize && blend_roi.w > 0) ? blend_roi.w : cw;
  const int bh = (minimize && blend_roi.h > 0) ? blend_roi.h : ch;
  const int bx0 = (minimize && blend_roi.w > 0) ? blend_roi.x : 0;
  const int by0 = (minimize && blend_roi.h > 0) ? blend_roi.y : 0;

  /* One-hot N-channel mask of the (cropped) seam: split_to_channels
   * (controlMasksN.cpp:221-232) converted to float (cudaPanoN.inl:192-194). */
  float* mask = (float*)malloc((size_t)bw * bh * n * sizeof(float));
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) {
      const int lbl = padded[(size_t)(by0 + y) * cw + (bx0 + x)];
      for (int i = 0; i < n; ++i) mask[((size_t)y * bw + x) * n + i] = (lbl == i) ? 1.0f : 0.0f;
    }
  const size_t fpx = (size_t)bw * bh * c * batch;
  float* full[HMO_MAX_IMAGES];
  for (int i = 0; i < n; ++i) full[i] = (float*)calloc(fpx, sizeof(float));
  float* out = (float*)calloc(fpx, sizeof(float));
  int rc = 0;

  if (use_min) {
    /* cudaPanoN.inl:545-641 */
    memset(canvas, 0, cpx * c * batch * ces);
    hmo_remap_hard_seam_n_t(n, inputs, pe, cfg->src_w, cfg->src_h, c, cfg->map_x, cfg->map_y, cfg->pos_x, cfg->pos_y,
                            cfg->map_w, cfg->map_h, padded, canvas, cw, ch, batch);
    for (int i = 0; i < n; ++i) {
      rect_t img = {cfg->pos_x[i], cfg->pos_y[i], cfg->map_w[i], cfg->map_h[i]};
      rect_t inter = rect_and(img, blend_roi);
      rect_t roi = {0, 0, 0, 0};
      if (inter.w > 0 && inter.h > 0) {
        roi.x = inter.x - cfg->pos_x[i];
        roi.y = inter.y - cfg->pos_y[i];
        roi.w = inter.w;
        roi.h = inter.h;
      }
      rc = hmo_remap_offset_t(inputs[i], pe, cfg->src_w[i], cfg->src_h[i], c, full[i], HMO_F32, bw, bh, c,
                              cfg->map_x[i], cfg->map_y[i], cfg->map_w[i], cfg->map_h[i], cfg->pos_x[i] - blend_roi.x,
                              cfg->pos_y[i] - blend_roi.y, roi.x, roi.y, roi.w, roi.h, 0, batch);
      if (rc) break;
    }
    if (!rc) rc = hmo_blend_n(n, (const float* const*)full, mask, out, bw, bh, c, cfg->levels, batch, dump);
    if (!rc)
      hmo_copy_roi_t(out, HMO_F32, bw, bh, c, write_roi.w, write_roi.h, write_roi.x - blend_roi.x,
                     write_roi.y - blend_roi.y, canvas, pe, cw, ch, c, write_roi.x, write_roi.y, batch);
  } else {
    /* cudaPanoN.inl:644-684 (the canvas is not cleared on this path). */
    for (int i = 0; i < n; ++i)
      hmo_remap_offset_t(inputs[i], pe, cfg->src_w[i], cfg->src_h[i], c, full[i], HMO_F32, bw, bh, c, cfg->map_x[i],
                         cfg->map_y[i], cfg->map_w[i], cfg->map_h[i], cfg->pos_x[i], cfg->pos_y[i], 0, 0, -1, -1, 0,
                         batch);
    rc = hmo_blend_n(n, (const float* const*)full, mask, out, bw, bh, c, cfg->levels, batch, dump);
    if (!rc) hmo_copy_roi_t(out, HMO_F32, bw, bh, c, bw, bh, 0, 0, canvas, pe, cw, ch, c, 0, 0, batch);
  }
  for (int i = 0; i < n; ++i) free(full[i]);
  free(out);
  free(mask);
  free(padded);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* 3-image pano                                                              */
/* ------------------------------------------------------------------------- */

/* ControlMasks3::canvas_width/height (src/pano/controlMasks3.cpp:303-317). */
int hmo_pano3_canvas_size(const hmo_pano3_cfg* cfg, int* canvas_w, int* canvas_h) {
  int w = 0, h = 0;
  for (int i = 0; i < 3; ++i) {
    w = IMAX(w, cfg->pos_x[i] + cfg->map_w[i]);
    h = IMAX(h, cfg->pos_y[i] + cfg->map_h[i]);
  }
  *canvas_w = w;
  *canvas_h = h;
  return 0;
}

/* FusedRemapToFullKernel3 (cudaFusedKernels3.cu:25-119): over the full
 * buffers' extent, each image's box writes only mapped, in-range pixels. */
static void fused_remap_to_full3(const hmo_pano3_cfg* cfg, const void* const* in, int pe, float* const* full, int fw,
                                 int fh) {
  const int c = cfg->channels;
  HMO_PARALLEL_ROWS
  for (int b = 0; b < cfg->batch; ++b)
    for (int y = 0; y < fh; ++y)
      for (int x = 0; x < fw; ++x)
        for (int i = 0; i < 3; ++i) {
          const int rx = x - cfg->pos_x[i], ry = y - cfg->pos_y[i];
          if (rx < 0 || rx >= cfg->map_w[i] || ry < 0 || ry >= cfg->map_h[i]) continue;
          const size_t mi = (size_t)ry * cfg->map_w[i] + rx;
          const int sx = cfg->map_x[i][mi], sy = cfg->map_y[i][mi];
          if (sx == HMO_UNMAPPED || !(sx < cfg->src_w[i] && sy < cfg->src_h[i])) continue;
          float v[4];
          load_any(in[i], pe, pix_index(b, sx, sy, cfg->src_w[i], cfg->src_h[i], c), c, v);
          store_any(full[i], HMO_F32, pix_index(b, x, y, fw, fh, c), c, v, pe);
        }
}

/* FusedRemapHardSeam3 (cudaFusedKernels3.cu:124-217): seam indexed with the
 * canvas width; unmapped / out-of-range pixels are left untouched. */
static void fused_hard_seam3(const hmo_pano3_cfg* cfg, const void* const* in, int pe, void* canvas, int cw, int ch) {
  const int c = cfg->channels;
  const size_t es = et_size(pe);
  HMO_PARALLEL_ROWS
  for (int b = 0; b < cfg->batch; ++b)
    for (int y = 0; y < ch; ++y)
      for (int x = 0; x < cw; ++x) {
        const int i = cfg->seam[(size_t)y * cw + x];
        if (i > 2) continue;
        const int rx = x - cfg->pos_x[i], ry = y - cfg->pos_y[i];
        if (rx < 0 || rx >= cfg->map_w[i] || ry < 0 || ry >= cfg->map_h[i]) continue;
        const size_t mi = (size_t)ry * cfg->map_w[i] + rx;
        const int sx = cfg->map_x[i][mi], sy = cfg->map_y[i][mi];
        if (sx == HMO_UNMAPPED || !(sx < cfg->src_w[i] && sy < cfg->src_h[i])) continue;
        memcpy((uint8_t*)canvas + pix_index(b, x, y, cw, ch, c) * es,
               (const uint8_t*)in[i] + pix_index(b, sx, sy, cfg->src_w[i], cfg->src_h[i], c) * es, (size_t)c * es);
      }
}

int hmo_pano3_process(const hmo_pano3_cfg* cfg, const void* const* inputs, void* canvas, hmo_dump* dump) {
  const int c = cfg->channels, batch = cfg->batch;
  const int pe = cfg->pipeline_f32 ? HMO_F32 : HMO_U8; /* element type of inputs and canvas */
  if ((c != 3 && c != 4) || batch < 1 || cfg->levels < 0) return -1;
  int cw, ch;
  hmo_pano3_canvas_size(cfg, &cw, &ch);
  const int hard = cfg->levels == 0;
  /* The hard seam indexes the seam with the canvas width (cudaFusedKernels3.cu:151,
   * cudaRemap.cu:233); the soft mask is cropped to canvas_h rows
   * (CanvasManager3::convertMaskMat, canvasManager3.cpp:107-128). */
  if (hard && (cfg->seam_w != cw || cfg->seam_h < ch)) return -1;
  if (!hard && cfg->seam_h < ch) return -1;
  if (c == 4) memset(canvas, 0, (size_t)cw * ch * c * batch * et_size(pe)); /* cudaPano3_optimized.inl:36-39 */
  if (hard) {
    if (cfg->fused) {
      fused_hard_seam3(cfg, inputs, pe, canvas, cw, ch);
    } else {
      for (int i = 0; i < 3; ++i) /* cudaPano3.inl: dest-map remaps in image order */
        hmo_remap_dest_map_t(inputs[i], pe, cfg->src_w[i], cfg->src_h[i], c, canvas, cw, ch, cfg->map_x[i],
                             cfg->map_y[i], cfg->map_w[i], cfg->map_h[i], cfg->pos_x[i], cfg->pos_y[i], i, cfg->seam,
                             batch);
    }
    return 0;
  }
  const int bw = cfg->seam_w, bh = ch;
  float* mask = (float*)malloc((size_t)bw * bh * 3 * sizeof(float));
  for (int y = 0; y < bh; ++y)
    for (int x = 0; x < bw; ++x) {
      const int lbl = cfg->seam[(size_t)y * cfg->seam_w + x];
      for (int i = 0; i < 3; ++i) mask[((size_t)y * bw + x) * 3 + i] = (lbl == i) ? 1.0f : 0.0f;
    }
  const size_t fpx = (size_t)bw * bh * c * batch;
  float* full[3];
  for (int i = 0; i < 3; ++i) full[i] = (float*)calloc(fpx, sizeof(float));
  if (cfg->fused) {
    fused_remap_to_full3(cfg, inputs, pe, full, bw, bh);
  } else {
    /* cudaPano3.inl:210-305: remap i onto the canvas, then copy its box into full i. */
    for (int i = 0; i < 3; ++i) {
      hmo_remap_offset_t(inputs[i], pe, cfg->src_w[i], cfg->src_h[i], c, canvas, pe, cw, ch, c, cfg->map_x[i],
                         cfg->map_y[i], cfg->map_w[i], cfg->map_h[i], cfg->pos_x[i], cfg->pos_y[i], 0, 0, -1, -1, 0,
                         batch);
      memset(full[i], 0, fpx * sizeof(float));
      hmo_copy_roi_t(canvas, pe, cw, ch, c, cfg->map_w[i], cfg->map_h[i], cfg->pos_x[i], cfg->pos_y[i], full[i],
                     HMO_F32, bw, bh, c, cfg->pos_x[i], cfg->pos_y[i], batch);
    }
  }
  float* out = (float*)calloc(fpx, sizeof(float));
  int rc = hmo_blend3(full[0], full[1], full[2], mask, out, bw, bh, c, cfg->levels, batch, dump);
  if (!rc) hmo_copy_roi_t(out, HMO_F32, bw, bh, c, bw, bh, 0, 0, canvas, pe, cw, ch, c, 0, 0, batch);
  for (int i = 0; i < 3; ++i) free(full[i]);
  free(out);
  free(mask);
  return rc;
}
