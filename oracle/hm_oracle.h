/*
 * hm_oracle.h -- CPU restatement of hm-cupano's single-frame stitch path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (hm-cupano_amd/) never links or calls it.
 *
 * Every function restates one reference function, in the reference's own
 * order of floating-point operations, and cites the file:line it follows
 * (paths relative to the hm-cupano source tree).  It is compiled with
 * -ffp-contract=off so that no multiply-add is fused, exactly like the HIP
 * kernels it checks.
 *
 * Conventions
 *   - pipeline images are uint8, tightly packed  B x H x W x C  (C = 3 or 4);
 *   - compute images are float, tightly packed   B x H x W x C;
 *   - blend masks are float, unbatched           H x W x M  (M = 1, 3 or N);
 *   - all functions return 0 on success, a negative value on invalid input.
 */
#ifndef HM_ORACLE_H
#define HM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HMO_UNMAPPED 65535

/* Element types of pipeline images (the reference's uchar, float and __half
 * vectors). */
#define HMO_U8 0
#define HMO_F32 1
#define HMO_F16 2

/* Per-level pyramid dump (gauss/mask/lap/blend/reconstruct) -- mirrors the
 * layout written by CudaStitchPano::dump_soft_blend_pyramid
 * (src/pano/cudaPano.inl:420-458).  Each pointer array has `levels` entries;
 * a NULL array or entry is skipped.  gauss/lap are indexed [image][level]. */
#define HMO_MAX_IMAGES 8
#define HMO_MAX_LEVELS 16
typedef struct hmo_dump {
  int levels;                               /* filled in by the oracle */
  int widths[HMO_MAX_LEVELS];
  int heights[HMO_MAX_LEVELS];
  float* gauss[HMO_MAX_IMAGES][HMO_MAX_LEVELS];
  float* lap[HMO_MAX_IMAGES][HMO_MAX_LEVELS];
  float* mask[HMO_MAX_LEVELS];
  float* blend[HMO_MAX_LEVELS];
  float* reconstruct[HMO_MAX_LEVELS];
} hmo_dump;

/* Threads of the per-pixel stages (OpenMP over rows); n <= 0 keeps the
 * current setting.  Returns the number now in use.  Results do not depend on
 * it: every output pixel is computed by one thread from its own inputs. */
int hmo_set_threads(int n);

/* ---- casts: src/cuda/cudaUtils.cuh:18-28 (truncating round_to_uchar) ---- */
uint8_t hmo_round_to_uchar(float x);

/* ---- pyramid level sizes ----
 * clamp_small_sides = 0: 2-image rule (src/cuda/cudaBlend.cu:860-866)
 * clamp_small_sides = 1: N/3-image rule, stop when a side < 2
 *                        (src/cuda/cudaBlendN.cu:675-691). Returns levels. */
int hmo_pyramid_sizes(int w, int h, int levels, int clamp_small_sides, int* widths, int* heights);

/* IEEE binary16 conversions: exact widening, round-to-nearest-even narrowing. */
float hmo_f16_to_f32(uint16_t h);
uint16_t hmo_f32_to_f16(float f);

/* ---- remap family: src/cuda/cudaRemap.cu ----
 * The _t forms take any element type (HMO_U8 / HMO_F32 / HMO_F16) and 1, 3
 * or 4 channels; the others are their uint8-source shorthands. */
int hmo_remap_offset_t(const void* src, int src_et, int src_w, int src_h, int src_c, void* dst, int dst_et,
                       int dst_w, int dst_h, int dst_c, const uint16_t* map_x, const uint16_t* map_y, int remap_w,
                       int remap_h, int off_x, int off_y, int roi_x, int roi_y, int roi_w, int roi_h,
                       int no_unmapped_write, int batch);
int hmo_remap_dest_map_t(const void* src, int et, int src_w, int src_h, int c, void* dst, int dst_w, int dst_h,
                         const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h, int off_x, int off_y,
                         int image_index, const uint8_t* dest_map, int batch);
int hmo_remap_hard_seam_n_t(int n, const void* const* srcs, int et, const int* src_w, const int* src_h, int c,
                            const uint16_t* const* map_x, const uint16_t* const* map_y, const int* off_x,
                            const int* off_y, const int* map_w, const int* map_h, const uint8_t* dest_map, void* dst,
                            int dst_w, int dst_h, int batch);
int hmo_copy_roi_t(const void* src, int src_et, int src_w, int src_h, int src_c, int region_w, int region_h,
                   int src_roi_x, int src_roi_y, void* dst, int dst_et, int dst_w, int dst_h, int dst_c, int off_x,
                   int off_y, int batch);
/* BatchedRemapKernelExOffset / ...ROI (cudaRemap.cu:61-198, host :326-390).
 * roi_w < 0 selects the whole map (non-ROI launcher). dst_is_float selects a
 * float destination (perform_cast uchar->float), else uint8. */
int hmo_remap_offset(const uint8_t* src, int src_w, int src_h, int src_c, void* dst, int dst_is_float, int dst_w,
                     int dst_h, int dst_c, const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h,
                     int off_x, int off_y, int roi_x, int roi_y, int roi_w, int roi_h, int no_unmapped_write,
                     int batch);
/* BatchedRemapKernelExOffsetWithDestMap (cudaRemap.cu:204-246). */
int hmo_remap_dest_map(const uint8_t* src, int src_w, int src_h, int c, uint8_t* dst, int dst_w, int dst_h,
                       const uint16_t* map_x, const uint16_t* map_y, int remap_w, int remap_h, int off_x, int off_y,
                       int image_index, const uint8_t* dest_map, int batch);
/* BatchedRemapHardSeamKernelN (cudaRemap.cu:252-294). */
int hmo_remap_hard_seam_n(int n, const uint8_t* const* srcs, const int* src_w, const int* src_h, int c,
                          const uint16_t* const* map_x, const uint16_t* const* map_y, const int* off_x,
                          const int* off_y, const int* map_w, const int* map_h, const uint8_t* dest_map, uint8_t* dst,
                          int dst_w, int dst_h, int batch);

/* ---- ROI copy: src/cuda/cudaMakeFull.cu:40-69 (copyRoiKernelBatched) ---- */
int hmo_copy_roi(const void* src, int src_is_float, int src_w, int src_h, int src_c, int region_w, int region_h,
                 int src_roi_x, int src_roi_y, void* dst, int dst_is_float, int dst_w, int dst_h, int dst_c,
                 int off_x, int off_y, int batch);

/* ---- single pyramid stages (float B x H x W x C images) ---- */
/* FusedBatchedDownsampleKernel, one image (src/cuda/cudaBlend.cu:46-144). */
void hmo_downsample(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int c, int batch);
/* BatchedDownsampleKernelMask / ...MaskN (cudaBlend.cu:268-296, cudaBlendN.cu:19-48). */
void hmo_downsample_mask(const float* in, int in_w, int in_h, float* out, int out_w, int out_h, int m);
/* BatchedComputeLaplacianKernel (cudaBlend.cu:356-456). */
void hmo_laplacian(const float* high, int hw, int hh, const float* low, int lw, int lh, float* lap, int c, int batch);
/* BatchedReconstructKernel (cudaBlend.cu:581-672). */
void hmo_reconstruct(const float* low, int lw, int lh, const float* lap, int hw, int hh, float* out, int c, int batch);

/* ---- Laplacian blends ---- */
/* cudaBatchedLaplacianBlendWithContext (src/cuda/cudaBlend.cu:847-1049). */
int hmo_blend2(const float* img1, const float* img2, const float* mask, float* out, int w, int h, int c, int levels,
               int batch, hmo_dump* dump);
/* cudaBatchedLaplacianBlendWithContextN (src/cuda/cudaBlendN.cu:649-828). */
int hmo_blend_n(int n, const float* const* imgs, const float* mask, float* out, int w, int h, int c, int levels,
                int batch, hmo_dump* dump);
/* cudaBatchedLaplacianBlendWithContext3 (src/cuda/cudaBlend3.cu:881-1206). */
int hmo_blend3(const float* img1, const float* img2, const float* img3, const float* mask, float* out, int w, int h,
               int c, int levels, int batch, hmo_dump* dump);

/* ---- seam padding: CanvasManager::convertMaskMat (src/pano/canvasManager.cpp:63-86,
 * src/pano/canvasManagerN.h:31-51) = cv::copyMakeBorder(mask, 0, ch-sh, 0,
 * cw-sw, BORDER_REPLICATE), i.e. out[y][x] = seam[min(y,sh-1)][min(x,sw-1)]
 * (right/bottom only).  out is cw x ch; -1 when the seam is larger than the
 * canvas (the reference asserts) or empty. ---- */
int hmo_pad_replicate(const uint8_t* seam, int sw, int sh, uint8_t* out, int cw, int ch);

/* ---- 2-image pano: CudaStitchPano ctor + process_impl ----
 * (src/pano/cudaPano.inl:19-94, 167-418; canvasManager.cpp:16-101) */
typedef struct hmo_pano2_cfg {
  int batch, levels, channels, minimize_blend;
  int src_w[2], src_h[2];
  int map_w[2], map_h[2];
  const uint16_t* map_x[2];
  const uint16_t* map_y[2];
  int pos_x[2], pos_y[2];    /* normalized integer positions */
  const uint8_t* seam;       /* post-load seam (1 = image 1), seam_w x seam_h */
  int seam_w, seam_h;
  int pipeline_f32;           /* 1: float pipeline images and canvas (CudaStitchPano*<float4|float3, ...>) */
} hmo_pano2_cfg;
int hmo_pano2_canvas_size(const hmo_pano2_cfg* cfg, int* canvas_w, int* canvas_h);
/* in1/in2/canvas: uint8, or float when cfg->pipeline_f32 */
int hmo_pano2_process(const hmo_pano2_cfg* cfg, const void* in1, const void* in2, void* canvas, hmo_dump* dump);

/* ---- N-image pano: CudaStitchPanoN ctor + process ----
 * (src/pano/cudaPanoN.inl:12-370, 489-685; canvasManagerN.h:13-79) */
typedef struct hmo_panoN_cfg {
  int n, batch, levels, channels, minimize_blend;
  int src_w[HMO_MAX_IMAGES], src_h[HMO_MAX_IMAGES];
  int map_w[HMO_MAX_IMAGES], map_h[HMO_MAX_IMAGES];
  const uint16_t* map_x[HMO_MAX_IMAGES];
  const uint16_t* map_y[HMO_MAX_IMAGES];
  int pos_x[HMO_MAX_IMAGES], pos_y[HMO_MAX_IMAGES];
  const uint8_t* seam;       /* labels 0..n-1, seam_w x seam_h */
  int seam_w, seam_h;
  int pipeline_f32;           /* 1: float pipeline images and canvas (CudaStitchPano*<float4|float3, ...>) */
} hmo_panoN_cfg;
int hmo_panoN_canvas_size(const hmo_panoN_cfg* cfg, int* canvas_w, int* canvas_h);
/* ROIs chosen by the ctor (for tests): write/blend ROI as x,y,w,h. */
int hmo_panoN_rois(const hmo_panoN_cfg* cfg, int* write_roi, int* blend_roi);
int hmo_panoN_process(const hmo_panoN_cfg* cfg, const void* const* inputs, void* canvas, hmo_dump* dump);

/* ---- 3-image pano: CudaStitchPano3 ctor + process ----
 * (src/pano/cudaPano3.inl:19-110, 183-422; cudaPano3_optimized.inl:19-163;
 *  canvasManager3.cpp:13-130; cudaFusedKernels3.cu:25-217).
 * The canvas is in/out: a 3-channel pipeline does not clear it. */
typedef struct hmo_pano3_cfg {
  int batch, levels, channels, fused;
  int src_w[3], src_h[3];
  int map_w[3], map_h[3];
  const uint16_t* map_x[3];
  const uint16_t* map_y[3];
  int pos_x[3], pos_y[3];
  const uint8_t* seam;       /* labels 0..2, seam_w x seam_h */
  int seam_w, seam_h;
  int pipeline_f32;           /* 1: float pipeline images and canvas (CudaStitchPano*<float4|float3, ...>) */
} hmo_pano3_cfg;
int hmo_pano3_canvas_size(const hmo_pano3_cfg* cfg, int* canvas_w, int* canvas_h);
int hmo_pano3_process(const hmo_pano3_cfg* cfg, const void* const* inputs, void* canvas, hmo_dump* dump);

#ifdef __cplusplus
}
#endif
#endif
