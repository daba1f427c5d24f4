/* hm_files.h -- the on-disk control data of a rig and the raster files of a
 * pyramid dump, as plain C.  Host code only: nothing here touches a GPU.
 *
 * A rig directory holds what Hugin's nona and enblend/multiblend write
 * (src/pano/controlMasks.cpp:273-326):
 *   mapping_%04d_x.tif, mapping_%04d_y.tif  uint16 remap tables, one sample
 *                                            per pixel (65535 = unmapped)
 *   mapping_%04d.tif                         the warped image; only its tags
 *                                            are read: XPOSITION x XRESOLUTION,
 *                                            YPOSITION x YRESOLUTION give the
 *                                            placement on the canvas
 *   seam_file.png                            the seam: an 8-bit paletted PNG
 *                                            (palette indices) or grayscale
 *
 * The decoders take baseline TIFF (little or big endian; strips or tiles;
 * chunky or planar; uncompressed, LZW, Deflate or PackBits; predictor 1, 2
 * or 3; 8/16/32-bit integer or 32/64-bit float samples) and PNG (every
 * filter, Adam7 interlace; paletted 8-bit, grayscale 1/2/4/8/16-bit,
 * grayscale + alpha).  Colour seam PNGs (RGB, RGBA, paletted below 8 bits)
 * are refused with HM_FILE_UNSUPPORTED: convert them to paletted or
 * grayscale first.
 *
 * Every function returns an hmFileStatus; hmFilesLastError() gives the text
 * of the calling thread's last failure. */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum hmFileStatus {
  HM_FILE_OK = 0,
  HM_FILE_NOT_FOUND = 1,        /* the file does not exist or cannot be opened */
  HM_FILE_BAD_FORMAT = 2,       /* not a TIFF / PNG, truncated or corrupt */
  HM_FILE_UNSUPPORTED = 3,      /* a valid file in a form the decoder does not take */
  HM_FILE_WRONG_TYPE = 4,       /* decoded, but not the pixel type the caller needs */
  HM_FILE_SEAM_CLASSES = 5,     /* seam label count differs from the image count */
  HM_FILE_MISMATCH = 6,         /* a rig's map_x and map_y differ in size */
  HM_FILE_WRITE_FAILED = 7,     /* a file or directory could not be written */
  HM_FILE_INVALID_ARGUMENT = 8
} hmFileStatus;

const char* hmFilesLastError(void);

/* ------------------------------------------------------------------------- */
/* Rig control directories                                                   */
/* ------------------------------------------------------------------------- */
enum {
  HM_CONTROL_FILES_MAX = 8,
  /* kind: which loader's rules apply */
  HM_CONTROL_TWO = 2,   /* ControlMasks::load (src/pano/controlMasks.cpp:273-326) */
  HM_CONTROL_THREE = 3, /* ControlMasks3::load (src/pano/controlMasks3.cpp:228-295) */
  HM_CONTROL_N = 0      /* ControlMasksN::load (src/pano/controlMasksN.cpp:125-193) */
};

typedef struct hmControlFiles hmControlFiles;

/* Host views into a loaded rig; valid until hmControlFilesDestroy. */
typedef struct hmControlFilesInfo {
  int n_images;
  uint32_t map_w[HM_CONTROL_FILES_MAX], map_h[HM_CONTROL_FILES_MAX];
  const uint16_t* map_x[HM_CONTROL_FILES_MAX]; /* row-major, map_w x map_h */
  const uint16_t* map_y[HM_CONTROL_FILES_MAX];
  /* Placements as the reference's SpatialTiff: position tag x resolution tag
   * in float, then the minimum over the images subtracted.  The pano classes
   * take them truncated toward zero (cv::Point of a float). */
  float xpos[HM_CONTROL_FILES_MAX], ypos[HM_CONTROL_FILES_MAX];
  /* Bit i: mapping_%04d.tif of image i was opened (a missing one gives a
   * zero placement, as in the reference, which only prints an error). */
  uint32_t position_files_found;
  uint32_t seam_w, seam_h;
  /* The seam as the reference's loader leaves it:
   *   TWO:   min value -> 1, max value -> 0 (1 selects image 1)
   *   THREE: the palette indices
   *   N:     labels 0..N-1 (sorted distinct values mapped to their rank) */
  const uint8_t* seam;
  int seam_paletted; /* 1: palette indices, 0: grayscale values */
} hmControlFilesInfo;

/* Loads dir/mapping_*.tif and dir/seam_file.png under the rules of `kind`.
 *   TWO:   images 0 and 1; seam paletted 8-bit or grayscale, then inverted
 *   THREE: images 0..2; the seam must be 8-bit paletted (the reference
 *          throws otherwise, controlMasks3.cpp:92-186)
 *   N:     images 0..n_images-1 (1 <= n_images <= 8); paletted or
 *          grayscale; exactly n_images distinct values (else
 *          HM_FILE_SEAM_CLASSES)
 * Maps must be one-sample uint16 (HM_FILE_WRONG_TYPE otherwise). */
int hmControlFilesLoad(const char* dir, int kind, int n_images, hmControlFiles** out);
int hmControlFilesGetInfo(const hmControlFiles* files, hmControlFilesInfo* info);
void hmControlFilesDestroy(hmControlFiles* files);

/* ------------------------------------------------------------------------- */
/* Raster files                                                              */
/* ------------------------------------------------------------------------- */
enum {
  HM_SAMPLE_UINT = 1,
  HM_SAMPLE_INT = 2,
  HM_SAMPLE_FLOAT = 3
};

/* An interleaved host raster (rows packed, host byte order). */
typedef struct hmRaster {
  void* data;
  uint32_t width, height, channels;
  uint32_t bits;          /* per sample: 8, 16, 32 or 64 */
  uint32_t sample_format; /* HM_SAMPLE_* */
  size_t bytes;
} hmRaster;

typedef struct hmTiffTags {
  int has_resolution; /* both XRESOLUTION and YRESOLUTION present */
  int has_xres, has_yres, has_xpos, has_ypos;
  float xres, yres, xpos, ypos; /* rationals as float(num / den) */
  uint32_t width, height, channels, bits, sample_format, compression, predictor, photometric;
} hmTiffTags;

/* Reads the first image of a TIFF into a library-allocated raster (release
 * with hmRasterFree).  opencv_order: for 3 and 4 samples, swap samples 0 and
 * 2 (OpenCV's BGR(A) memory order of an RGB(A) file). */
int hmTiffRead(const char* path, int opencv_order, hmRaster* out);
/* Reads only the first IFD's tags. */
int hmTiffReadTags(const char* path, hmTiffTags* tags);
/* Writes an uncompressed little-endian TIFF (1..8 samples; 3 or 4 samples as
 * RGB(A), swapped like OpenCV's writer when opencv_order is set).  tags may be
 * NULL; its has_* fields select XRESOLUTION/YRESOLUTION/XPOSITION/YPOSITION. */
int hmTiffWrite(const char* path, const hmRaster* raster, const hmTiffTags* tags, int opencv_order);
/* Reads a seam PNG as one 8-bit channel: palette indices of an 8-bit
 * paletted file (*paletted = 1) or the gray value of a grayscale one (16-bit
 * keeps the high byte; 1/2/4-bit values are scaled to 0..255; gray + alpha
 * drops the alpha). */
int hmPngReadSeam(const char* path, hmRaster* out, int* paletted);
void hmRasterFree(hmRaster* raster);

#ifdef __cplusplus
}
#endif
