/*
 * hm_stitch.h -- C ABI of the MI355X (gfx950) stitch library libhm_stitch.so.
 *
 * This is the drop-in boundary for hm-cupano's single-frame stitch path.  It
 * replaces two tiers of the reference (paths relative to the hm-cupano tree):
 *
 *   tier 1, kernel host functions  src/cuda/cudaRemap.h, cudaMakeFull.h,
 *           cudaBlend.h, cudaBlendN.h, cudaBlend3.h   -> hmRemap*, hmCopyRoi,
 *           hmMakeFull, hmBlend*
 *   tier 2, pano classes           src/pano/cudaPano.h, cudaPanoN.h,
 *           cudaPano3.h                                -> hmStitchPano*
 *
 * The C++ classes with the reference names (hm::pano::cuda::CudaStitchPano,
 * CudaStitchPanoN, CudaStitchPano3, hm::CudaMat, CudaStatus ...) are
 * header-only wrappers over this ABI (hm-cupano_amd/include/cupano/).
 *
 * Conventions
 *   - every function returns an int holding a hipError_t value (0 = success);
 *     launchers are asynchronous on the given stream, like the reference's;
 *   - device memory is passed as plain pointers (hmSurface for pitched
 *     images); host arrays (maps, seams) are plain host pointers;
 *   - pixel types: pipeline images are uint8 BGRA (4 channels) or BGR
 *     (3 channels); compute images are float with the same channel count.
 */
#ifndef HM_STITCH_H
#define HM_STITCH_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Pitched, batch-major device image: pixel (b,x,y) lives at
 * d_ptr + (b*height + y)*pitch + x*sizeof(pixel).
 * Replaces CudaSurface<T> (src/cuda/cudaTypes.h:195-201). */
typedef struct hmSurface {
  void* d_ptr;
  uint32_t width;
  uint32_t height;
  uint32_t pitch; /* bytes */
} hmSurface;

/* Pixel types of the template instantiations the reference ships
 * (src/cuda/cudaRemap.cu:507-554, cudaMakeFull.cu:291-313). */
typedef enum hmPixelType {
  HM_PIX_U8C3 = 0,  /* uchar3 */
  HM_PIX_U8C4 = 1,  /* uchar4 */
  HM_PIX_F32C3 = 2, /* float3 */
  HM_PIX_F32C4 = 3, /* float4 */
  HM_PIX_U8C1 = 4,  /* uchar1 */
  HM_PIX_F32C1 = 5, /* float1 */
  HM_PIX_F16C1 = 6, /* __half */
  HM_PIX_F16C3 = 7, /* half3 */
  HM_PIX_F16C4 = 8  /* half4 */
} hmPixelType;

/* ------------------------------------------------------------------------- */
/* Tier 1: remap and ROI copy                                                */
/* ------------------------------------------------------------------------- */

/* batched_remap_kernel_ex_offset (src/cuda/cudaRemap.h:41-54,
 * cudaRemap.cu:326-344).  src and dest have the same channel count; dest is
 * the source type, or float for a uint8 source (the reference instantiates
 * uchar3/4 -> uchar or float, float1/3/4 -> float).  The default pixel is
 * all-zero (T_in{}), as every pano class passes. */
int hmRemapExOffset(hmSurface src, hmPixelType src_type, hmSurface dest, hmPixelType dest_type,
                    const uint16_t* d_map_x, const uint16_t* d_map_y, int batch_size, int remap_w, int remap_h,
                    int offset_x, int offset_y, int no_unmapped_write, hipStream_t stream);

/* batched_remap_kernel_ex_offset_roi (src/cuda/cudaRemap.h:72-89,
 * cudaRemap.cu:347-390), including its ROI validation. */
int hmRemapExOffsetRoi(hmSurface src, hmPixelType src_type, hmSurface dest, hmPixelType dest_type,
                       const uint16_t* d_map_x, const uint16_t* d_map_y, int batch_size, int remap_w, int remap_h,
                       int offset_x, int offset_y, int roi_x, int roi_y, int roi_w, int roi_h,
                       int no_unmapped_write, hipStream_t stream);

/* batched_remap_kernel_ex_offset_with_dest_map (src/cuda/cudaRemap.h:91-105,
 * cudaRemap.cu:396-415).  dest_image_map is unpitched, dest.width wide.
 * Any pixel type (the reference: uchar1/3/4, float1/3/4, __half). */
int hmRemapExOffsetWithDestMap(hmSurface src, hmSurface dest, hmPixelType type, const uint16_t* d_map_x,
                               const uint16_t* d_map_y, int this_image_index, const uint8_t* d_dest_image_map,
                               int batch_size, int remap_w, int remap_h, int offset_x, int offset_y,
                               hipStream_t stream);

/* batched_remap_hard_seam_kernel_n (src/cuda/cudaRemap.h:117-128,
 * cudaRemap.cu:418-437).  The per-image arrays are HOST arrays (their
 * contents travel in the kernel arguments; no per-frame H2D copy). */
int hmRemapHardSeamN(int n_images, const hmSurface* inputs, const uint16_t* const* d_map_x,
                     const uint16_t* const* d_map_y, const int* offsets_xy, const int* sizes_wh,
                     const uint8_t* d_dest_image_map, hmSurface dest, hmPixelType type, int batch_size,
                     hipStream_t stream);

/* copy_roi_batched (src/cuda/cudaMakeFull.h:61-86, cudaMakeFull.cu:241-258):
 * typed ROI copy with perform_cast (src/cuda/cudaUtils.cuh:80-213): float or
 * half -> uint8 clamps and truncates, -> half rounds to nearest even, 3 -> 4
 * channels sets alpha 255, 4 -> 3 drops it; 1-channel types copy to 1 channel. */
int hmCopyRoi(hmSurface src, hmPixelType src_type, int region_w, int region_h, int src_roi_x, int src_roi_y,
              hmSurface dest, hmPixelType dest_type, int offset_x, int offset_y, int batch_size,
              hipStream_t stream);

/* simple_make_full_batch (src/cuda/cudaMakeFull.h:32-46,
 * cudaMakeFull.cu:173-214): zero the destination, then hmCopyRoi. */
int hmMakeFull(hmSurface src, hmPixelType src_type, int region_w, int region_h, int src_roi_x, int src_roi_y,
               int dest_offset_x, int dest_offset_y, int batch_size, hmSurface dest, hmPixelType dest_type,
               hipStream_t stream);

/* ------------------------------------------------------------------------- */
/* Tier 1: Laplacian blend contexts                                          */
/* ------------------------------------------------------------------------- */
typedef struct hmBlendContext hmBlendContext;

typedef enum hmBlendKind {
  HM_BLEND_TWO = 2,   /* CudaBatchLaplacianBlendContext      (src/cuda/cudaBlend.h:38-141)  */
  HM_BLEND_THREE = 3, /* CudaBatchLaplacianBlendContext3     (src/cuda/cudaBlend3.h:45-161) */
  HM_BLEND_N = 0      /* CudaBatchLaplacianBlendContextN<N>  (src/cuda/cudaBlendN.h:15-99)  */
} hmBlendKind;

typedef enum hmLevelKind {
  HM_LEVEL_GAUSS = 0,
  HM_LEVEL_MASK = 1,
  HM_LEVEL_LAPLACIAN = 2,
  HM_LEVEL_BLEND = 3,
  HM_LEVEL_RECONSTRUCT = 4
} hmLevelKind;

/* Context constructor (cudaBlend.h:52-65, cudaBlendN.h:15-99,
 * cudaBlend3.h:45-161).  All levels are allocated here, not on first use.
 * debug != 0 also keeps the Laplacian and blended pyramids for dumps. */
int hmBlendContextCreate(int kind, int n_images, int width, int height, int num_levels, int batch_size, int channels,
                         int debug, hmBlendContext** out);
/* Actual level count and sizes (the N/3-image rule may lower the count). */
int hmBlendContextLevels(const hmBlendContext* ctx, int* num_levels, int* widths, int* heights);
/* cudaBatchedLaplacianBlendWithContext (src/cuda/cudaBlend.h:237-245),
 * ...WithContextN (cudaBlendN.h:182-188), ...WithContext3 (cudaBlend3.h).
 * d_images is a HOST array of n device pointers (tight B x H x W x C float);
 * d_mask is H x W x M float (M = 1, 3 or N); d_output may alias d_images[0]. */
int hmBlendRun(hmBlendContext* ctx, const float* const* d_images, const float* d_mask, float* d_output,
               hipStream_t stream);
/* Device pointer of one pyramid level (for dumps); NULL when not kept. */
int hmBlendContextLevelPtr(const hmBlendContext* ctx, int what, int image, int level, const float** ptr);
void hmBlendContextDestroy(hmBlendContext* ctx);

/* Bytes of the device scratch that hmBlendRunNoAlloc carves its pyramid
 * levels from: one array per GPU, held from the first call on that GPU until
 * the process ends. */
#define HM_NOALLOC_SCRATCH_BYTES (128u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Equivalent to hmBlendContextCreate(kind, n_images,
 * width, height, num_levels, batch_size, channels, 0, &c) + hmBlendRun(c,
 * d_images, d_mask, d_output, stream) + hmBlendContextDestroy(c), same
 * kernels and same bits, but it calls no hipMalloc/hipFree: levels >= 1 and
 * the mask pyramid live in a static device array of HM_NOALLOC_SCRATCH_BYTES.
 * Returns hipErrorOutOfMemory when the pyramid does not fit it (a 2-image,
 * 4-channel blend of 1200 x 3300 at 6 levels, batch 1, needs about 69 MB).
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (other streams or threads) race.  Nothing in it needs clearing: every
 * level is written in full before it is read. */
int hmBlendRunNoAlloc(int kind, int n_images, int width, int height, int num_levels, int batch_size, int channels,
                      const float* const* d_images, const float* d_mask, float* d_output, hipStream_t stream);

/* ------------------------------------------------------------------------- */
/* Tier 2: pano classes                                                      */
/* ------------------------------------------------------------------------- */

/* In-memory ControlMasks (src/pano/controlMasks.h:30-66): uint16 CV_16U
 * remap tables (host, tight), the seam AFTER load_seam_mask's inversion
 * (controlMasks.cpp:225-251: 1 selects image 1), and the SpatialTiff
 * positions (already normalised, controlMasks.cpp:42-59). */
typedef struct hmControlMasks2 {
  const uint16_t* img1_col;
  const uint16_t* img1_row;
  int img1_w, img1_h;
  const uint16_t* img2_col;
  const uint16_t* img2_row;
  int img2_w, img2_h;
  const uint8_t* whole_seam_mask_image;
  int seam_w, seam_h;
  float xpos[2], ypos[2];
} hmControlMasks2;

typedef struct hmStitchPano hmStitchPano;

/* The pano constructors' `channels` argument: 3 or 4 for the uint8 pipeline
 * (CudaStitchPano*<uchar3|uchar4, float3|float4>), or-ed with
 * HM_PIPELINE_F32 for the float pipeline (CudaStitchPano*<float3|float4,
 * float3|float4>: float inputs and a float canvas, no truncation, as in
 * src/pano/cudaPano_test.cpp:155, cudaPanoN_test.cpp:154, cudaPano3_test.cpp:173). */
#define HM_PIPELINE_F32 0x100

/* CudaStitchPano<uchar4|uchar3, float4|float3> ctor (src/pano/cudaPano.h:71-77,
 * cudaPano.inl:19-94).  channels = 4 (uchar4/float4) or 3 (uchar3/float3).
 * Like the reference, a failing constructor still returns an object whose
 * status is sticky; *out is NULL only on allocation failure of the object. */
int hmStitchPanoCreate(int batch_size, int num_levels, const hmControlMasks2* masks, int quiet, int minimize_blend,
                       int max_output_width, int channels, hmStitchPano** out);
/* CudaStitchPano::process (cudaPano.h:95-99, cudaPano.inl:460-472). */
int hmStitchPanoProcess(hmStitchPano* pano, hmSurface input1, hmSurface input2, hmSurface canvas,
                        hipStream_t stream);
/* Shared by every pano object (2-, 3- and N-image):
 * status() / canvas_width() / canvas_height() / batch_size()
 * (cudaPano.h:79-93, cudaPanoN.h:84-95, cudaPano3.h:72-83); Destroy replaces
 * the destructor. */
int hmStitchPanoStatus(const hmStitchPano* pano);
int hmStitchPanoCanvasSize(const hmStitchPano* pano, int* width, int* height);
int hmStitchPanoBatchSize(const hmStitchPano* pano);
void hmStitchPanoDestroy(hmStitchPano* pano);

/* No counterpart in the reference (it allocates lazily in the first
 * process): allocate everything a process call needs now, for callers that
 * must not allocate later (a serving loop, graph capture).  Errors are
 * sticky. */
int hmStitchPanoPrepare(hmStitchPano* pano);
/* No counterpart in the reference: device memory held by the object (the
 * allocation_size accounting of src/cuda/cudaBlend.h:107, for the whole
 * pano), used for admission against a memory budget. */
int hmStitchPanoDeviceBytes(const hmStitchPano* pano, size_t* bytes);

/* Launch-shape knobs of a pano object.  No counterpart in the reference.
 * Every setting computes the same output; only the speed differs.
 *   remap_px   pixels per thread of the 2-image remap kernels: 2, 4 or 8;
 *   tail_base  soft seam with 2 to 4 images: levels tail_base..L-1 of the up
 *              pass run as one tile-fused launch (1..3; with 3 or 4 images
 *              when at most 5 levels are in it), or 0: one launch per level;
 *   overlap    1: the canvas outside the blend region is remapped on a second
 *              stream overlapping the pyramid, forked after the blend inputs;
 *              2 (2-image class): the same, forked at the start of the frame;
 *              3 (2-image class): forked after the down pass, beside the up
 *              pass's small levels;
 *              0: everything on one stream. */
typedef struct hmTuning {
  int remap_px;
  int tail_base;
  int overlap;
  int remap_xcd;  /* 1: 2-image remap tiles ordered as horizontal canvas bands, one per XCD */
  int remap_rows;  /* canvas rows per thread of the 2-image remaps: 1, 2 or 4 (taller tiles) */
  int canvas_grid; /* workgroups of the 2-image canvas remap on the second stream: 0 = one per tile,
                      else 256, 512, 1024 or 2048 walking the tiles (a smaller share of the GPU) */
  int fuse_down;   /* 1: the 2-image blend-input remap runs inside the first down-pass launch */
  int fuse_up;     /* 1: levels 1 and 0 of the up pass run as one launch (level 1's reconstruction
                      stays in LDS): the 2-image soft seam, and the N-image class (3..8 images, uint8,
                      levels >= 3) where no level-0 tile meets more than 4 images */
} hmTuning;
int hmStitchPanoGetTuning(const hmStitchPano* pano, hmTuning* tuning);
/* SetTuning and Autotune return hipErrorNotSupported while a server holds the
 * object (between hmServerAddRig and hmServerRemoveRig / hmServerDestroy):
 * tune before adding the rig. */
int hmStitchPanoSetTuning(hmStitchPano* pano, const hmTuning* tuning);
/* The project's autotuner: times every setting of the knobs that matter for
 * this object on the caller's frame (inputs as for the class's process;
 * fused: the 3-image path), `frames` frames each after one warm-up frame,
 * keeps the fastest, and returns it with its ms/frame.  The canvas is
 * written. */
int hmStitchPanoAutotune(hmStitchPano* pano, const hmSurface* inputs, int num_inputs, hmSurface canvas, int fused,
                         int frames, hipStream_t stream, hmTuning* best, float* best_ms_per_frame);

#define HM_MAX_IMAGES 8

/* In-memory ControlMasksN (src/pano/controlMasksN.h:13-38): per-image CV_16U
 * remap tables (host, tight), the indexed seam with labels 0..n-1 (as
 * ControlMasksN::load leaves it, controlMasksN.cpp:125-193), and the
 * normalised SpatialTiff positions. */
typedef struct hmControlMasksN {
  int n_images;
  const uint16_t* img_col[HM_MAX_IMAGES];
  const uint16_t* img_row[HM_MAX_IMAGES];
  int img_w[HM_MAX_IMAGES], img_h[HM_MAX_IMAGES];
  const uint8_t* whole_seam_mask_indexed;
  int seam_w, seam_h;
  float xpos[HM_MAX_IMAGES], ypos[HM_MAX_IMAGES];
} hmControlMasksN;

/* CudaStitchPanoN<uchar4|uchar3, float4|float3> ctor (src/pano/cudaPanoN.h:83,
 * cudaPanoN.inl:108-370), same argument order (minimize_blend before quiet).
 * Constructor errors are sticky, as for the 2-image class. */
int hmStitchPanoNCreate(int batch_size, int num_levels, const hmControlMasksN* masks, int minimize_blend, int quiet,
                        int channels, hmStitchPano** out);
/* CudaStitchPanoN::process (cudaPanoN.h:97-101, cudaPanoN.inl:489-685):
 * inputs[0..num_inputs) are the N pitched pipeline images.  Like the
 * reference, validation errors are returned without becoming sticky. */
int hmStitchPanoNProcess(hmStitchPano* pano, const hmSurface* inputs, int num_inputs, hmSurface canvas,
                         hipStream_t stream);

/* In-memory ControlMasks3 (src/pano/controlMasks3.h:12-43): img0..img2
 * col/row CV_16U tables as arrays, the single-channel seam with labels 0..2,
 * and the three normalised positions. */
typedef struct hmControlMasks3 {
  const uint16_t* img_col[3];
  const uint16_t* img_row[3];
  int img_w[3], img_h[3];
  const uint8_t* whole_seam_mask_image;
  int seam_w, seam_h;
  float xpos[3], ypos[3];
} hmControlMasks3;

/* CudaStitchPano3<uchar4|uchar3, float4|float3> ctor (src/pano/cudaPano3.h:70,
 * cudaPano3.inl:26-110).  Constructor errors are sticky. */
int hmStitchPano3Create(int batch_size, int num_levels, const hmControlMasks3* masks, int quiet, int channels,
                        hmStitchPano** out);
/* CudaStitchPano3::process (cudaPano3.h:101-107, cudaPano3.inl:404-422):
 * fused != 0 is the default process_optimized path
 * (cudaPano3_optimized.inl:19-163).  Errors are sticky on both paths.  A
 * 3-channel canvas is not cleared: pixels no image writes keep their value. */
int hmStitchPano3Process(hmStitchPano* pano, hmSurface input0, hmSurface input1, hmSurface input2, hmSurface canvas,
                         int fused, hipStream_t stream);

/* ------------------------------------------------------------------------- */
/* Served form: many rigs, one GPU                                           */
/* ------------------------------------------------------------------------- */
/* No counterpart in the reference, whose apps call process() once per frame
 * and synchronise (tests/stitch_two_videos.cpp:552-585).  For a frame loop
 * that serves many camera rigs: rigs (pano objects of any class) are
 * admitted against a device-memory budget and keep their state resident;
 * frames are submitted as requests and started by a scheduler thread on one
 * of n_streams streams as soon as a stream and the rig are free, so work of
 * different rigs overlaps on the device; each rig's pipeline is captured as a
 * hipGraph per (inputs, canvas) buffer set and replayed.  Output is
 * bit-identical to the pano's own process().  The caller keeps ownership of
 * the pano objects and of every buffer until its request has completed. */
typedef struct hmServer hmServer;

typedef struct hmServerResult {
  uint64_t tag;        /* the caller's tag of the request */
  int rig_id;
  int status;          /* hipError_t of the request */
  int64_t t_submit_ns; /* steady clock: hmServerSubmit */
  int64_t t_start_ns;  /* launched on a stream */
  int64_t t_done_ns;   /* completion observed */
} hmServerResult;

typedef struct hmServerStats {
  int n_rigs;
  size_t device_budget_bytes;
  size_t device_used_bytes;
  int queued;
  int in_flight;
  long graphs_captured;
  long graph_launches;
  long eager_launches;
  long batched_passes;       /* batched passes started (hmServerSetBatching) */
  long batched_frames;       /* requests those passes carried */
  long pass_graphs_captured; /* passes captured as hipGraphs (a composition seen twice) */
  long pass_graph_launches;  /* passes replayed from a captured graph */
} hmServerStats;

int hmServerCreate(int n_streams, size_t device_budget_bytes, int use_graphs, hmServer** out);
/* Prepares the pano (hmStitchPanoPrepare) and admits it if its device bytes
 * fit the remaining budget (else hipErrorOutOfMemory, not admitted).  A pano
 * is held by one server at a time (else hipErrorInvalidValue); while held, its
 * tuning is fixed and it cannot join an hmBatch (hipErrorNotSupported). */
int hmServerAddRig(hmServer* server, hmStitchPano* pano, int* rig_id);
/* Waits for the rig's requests, then stops serving it (the pano survives). */
int hmServerRemoveRig(hmServer* server, int rig_id);
/* Queues one frame; returns at once.  fused: the 3-image class's path. */
int hmServerSubmit(hmServer* server, int rig_id, const hmSurface* inputs, int num_inputs, hmSurface canvas, int fused,
                   uint64_t tag);
/* Queues several frames at once, in order (e.g. the frames of many rigs
 * captured at one tick): the scheduler sees all of them together, so with
 * batching on they can share a pass.  All or nothing: one invalid request
 * (unknown rig, wrong input count) queues none. */
typedef struct hmServerRequest {
  int rig_id;
  const hmSurface* inputs;
  int num_inputs;
  hmSurface canvas;
  int fused;
  uint64_t tag;
} hmServerRequest;
int hmServerSubmitMany(hmServer* server, const hmServerRequest* requests, int num_requests);
/* Moves up to max_results completed requests out (oldest first). */
int hmServerPoll(hmServer* server, hmServerResult* results, int max_results, int* num_results);
/* Blocks until nothing is queued or running; returns the first error seen. */
int hmServerWaitIdle(hmServer* server);
int hmServerGetStats(hmServer* server, hmServerStats* stats);
/* Continuous batching (no counterpart in the reference).  max_frames >= 2:
 * when a stream frees up, the requests that are ready -- the oldest whose
 * rig is idle, then up to max_frames - 1 more of other idle rigs, oldest
 * first -- run as one batched pass (the persistent work-list kernels of
 * hmBatch, over each request's recorded launches); a request ready alone
 * takes its rig's graph.  A pass whose composition of (rig, buffer set)
 * comes back is captured as a hipGraph on its second sighting and replayed.
 * Rigs whose pipeline the batched kernels do not carry (float pipelines,
 * debug mode) always run alone.  0 or 1: off (the default).  policy 0: any
 * ready requests share a pass (every class, canvas size, image and level
 * count); policy 1: only requests whose pipelines have the same structure
 * (class, image and level count, channels: the same kernel bodies in the
 * same order, canvas sizes may differ), so no frame's stage waits for a
 * longer pipeline's.  Output is bit-identical either way. */
int hmServerSetBatching(hmServer* server, int max_frames, int policy);
/* Drains the queue, then releases streams and graphs. */
void hmServerDestroy(hmServer* server);

/* -------------------------------------------------------------------------
 * Batched pass over mixed pano configurations.  No counterpart in the
 * reference (its apps call process() once per frame).  Several pano calls --
 * rigs of the 2-, 3- and N-image classes with different image counts, canvas
 * sizes and level counts, each call possibly a batch of frames -- run as one
 * pass: launch k of every call forms stage k, and each kernel body of a stage
 * is one launch of a persistent kernel whose workgroups walk the tile
 * work-list of every call's launch of that body (the bodies of a stage run
 * concurrently).  The
 * pass is built for the given buffers (inputs and canvases are read and
 * written where they are at create time, like a captured graph); each frame's
 * output is bit-identical to its own process().
 *
 * Carried: the uint8 pipelines (3 or 4 channels) of the 2-image class, the
 * 3-image class and the N-image class with 2 to 8 images (any levels).  A
 * float pipeline, a pano in debug mode or being served, or one pano in two
 * calls: hipErrorNotSupported / hipErrorInvalidValue at create time.
 * ------------------------------------------------------------------------- */
typedef struct hmBatch hmBatch;
typedef struct hmBatchCall {
  hmStitchPano* pano;
  const hmSurface* inputs; /* the pano's inputs (2, 3 or N surfaces) */
  int num_inputs;
  hmSurface canvas;
  int fused; /* the 3-image class's path (see hmStitchPano3Process) */
} hmBatchCall;
/* Prepares every pano, records each call's launches and builds the stages. */
int hmBatchCreate(const hmBatchCall* calls, int num_calls, hmBatch** out);
/* Workgroups per launch: 0 (the default) one per work unit; fewer makes the
 * grid persistent, each workgroup walking the list with that stride. */
int hmBatchSetGrid(hmBatch* batch, int workgroups);
/* Times the pass (passes runs after a warm-up) for one workgroup per work
 * unit and for persistent grids of 2, 4, 8 and 16 per compute unit, and keeps
 * the fastest. */
int hmBatchAutotune(hmBatch* batch, hipStream_t stream, int passes, int* best_grid, float* best_ms_per_pass);
/* stages: pipeline stages; launches: kernel launches of one pass; recorded:
 * the launches the calls make one by one; grid: workgroups per launch. */
int hmBatchGetInfo(const hmBatch* batch, int* stages, int* launches, int* recorded_launches, int* grid);
/* Runs the pass on `stream`; use_graph != 0 captures it as a hipGraph on the
 * first call and replays that graph afterwards. */
int hmBatchRun(hmBatch* batch, hipStream_t stream, int use_graph);
int hmBatchDestroy(hmBatch* batch);

/* ------------------------------------------------------------------------- */
/* Frame loop (host frames in, host frames out)                              */
/* ------------------------------------------------------------------------- */
/* tests/stitch_two_videos.cpp:552-585 (read -> upload -> process -> download
 * -> BGRA2BGR -> write) as a pipeline: `slots` pinned host slots, uploads on
 * one stream, the stitch on `stream`, downloads on another, so frame k+1's
 * upload, frame k's stitch and frame k-1's download overlap.  Frames cross
 * PCIe as packed BGR (3 bytes per pixel); a 4-channel pano gets them
 * expanded to BGRA (alpha 255) and its canvas packed back to BGR on the
 * device.  One canvas serves every frame, as in the reference's loop.
 *
 * source(user, frame, host_frames, pitches) fills, for each input i, batch x
 * in_h[i] rows of in_w[i] BGR pixels at host_frames[i] (row pitch
 * pitches[i]); it returns 1 for a frame, 0 at the end of the stream, < 0 on
 * error.  sink(user, frame, bgr, width, rows, pitch) receives each stitched
 * frame in order (batch x canvas-height rows; valid during the call only);
 * nonzero stops the loop.  A
 * callback's error ends the loop with hipErrorUnknown and its value in
 * stats->callback_error. */
typedef int (*hmFrameSource)(void* user, int64_t frame, void* const* host_frames, const uint32_t* pitches);
typedef int (*hmFrameSink)(void* user, int64_t frame, const void* bgr, uint32_t width, uint32_t rows,
                           uint32_t pitch);
enum {
  HM_TRANSFER_DMA = 0,    /* hipMemcpyAsync on a copy stream */
  HM_TRANSFER_KERNEL = 1, /* a copy kernel reading / writing the mapped pinned buffer */
  HM_TRANSFER_DEFAULT = HM_TRANSFER_DMA
};
typedef struct hmFrameLoopOptions {
  int slots;          /* pinned host slots in flight, 1..8 (0: 3) */
  int64_t max_frames; /* < 0: until the source ends */
  int upload;         /* HM_TRANSFER_* for host -> device */
  int download;       /* HM_TRANSFER_* for device -> host */
} hmFrameLoopOptions;
typedef struct hmFrameLoopStats {
  int64_t frames;                /* frames delivered to the sink */
  double seconds;                /* wall time of the loop */
  double upload_ms, stitch_ms, download_ms; /* mean per-frame stream times */
  uint64_t bytes_up_per_frame, bytes_down_per_frame;
  int callback_error;
  /* The pipeline once filled: the first frame reaches the sink after every slot
   * has been filled once by the source (first_frame_seconds from the start);
   * steady_frames = frames - 1 delivered after it, in steady_seconds. */
  double first_frame_seconds;
  double steady_seconds;
  int64_t steady_frames;
} hmFrameLoopStats;
/* options may be NULL (3 slots, until the source ends, default transfers). */
int hmFrameLoopRun(hmStitchPano* pano, int n_inputs, const uint32_t* in_w, const uint32_t* in_h,
                   const hmFrameLoopOptions* options, hmFrameSource source, hmFrameSink sink, void* user,
                   hipStream_t stream, hmFrameLoopStats* stats);

/* ------------------------------------------------------------------------- */
/* Pyramid inspection and launch timing (all pano classes)                   */
/* ------------------------------------------------------------------------- */
/* Keep the Laplacian/blend pyramids (call before the first process). */
int hmStitchPanoSetDebug(hmStitchPano* pano, int enable);
int hmStitchPanoLevels(const hmStitchPano* pano, int* num_levels, int* widths, int* heights);
/* Copy one level to host memory (synchronises the stream), the data that
 * dump_soft_blend_pyramid writes (cudaPano.inl:420-458).  Level-0 Gaussians
 * are converted from the resident uint8 form to float.  Complete in debug
 * mode; otherwise the N- and 3-image Gaussian levels are kept only around
 * each image's box (elsewhere they are zero and not stored). */
int hmStitchPanoReadLevel(hmStitchPano* pano, int what, int image, int level, void* host_dst, size_t bytes,
                          hipStream_t stream);
/* dump_soft_blend_pyramid (src/pano/cudaPano.inl:420-458): synchronises the
 * stream, then writes `directory`/metadata.txt ("num_levels=", "batch_size=",
 * "channels=", "level_k=WxH" lines) and one float32 TIFF per level:
 *   gauss<i>/level_k_batch_b.tiff  (i = 1..N)   mask/level_k.tiff
 *   lap<i>/level_k_batch_b.tiff    blend/level_k_batch_b.tiff
 *   reconstruct/level_k_batch_b.tiff
 * (two images: gauss1 gauss2 lap1 lap2, the reference's names), in
 * cv::imwrite's channel order.  The levels are those of the last processed
 * frame, re-evaluated densely from its resident level-0 inputs (works with
 * or without hmStitchPanoSetDebug).  hipErrorInvalidValue for a hard seam
 * (levels 0) or before the first process. */
int hmStitchPanoDumpPyramid(hmStitchPano* pano, const char* directory, hipStream_t stream);
/* Per-launch event timing on the process stream. */
int hmStitchPanoSetTiming(hmStitchPano* pano, int enable);
/* Writes "name total_ms count\n" lines accumulated since the last reset. */
int hmStitchPanoTimingReport(hmStitchPano* pano, char* buf, size_t len, int reset);

#ifdef __cplusplus
}
#endif
#endif /* HM_STITCH_H */
