// controlMasks.h -- the control data of a rig, in memory or loaded from a
// control directory.
//
// The fields of hm-cupano's ControlMasks (src/pano/controlMasks.h:30-66):
// img1_col/img1_row/img2_col/img2_row are the CV_16U remap tables,
// whole_seam_mask_image is the seam AFTER load_seam_mask's inversion
// (controlMasks.cpp:225-251: 1 selects image 1), positions are the normalised
// SpatialTiff placements.  cv::Mat is replaced by a minimal host image.
// load() reads the directory through hmControlFilesLoad (include/hm_files.h)
// under the same rules as the reference's loaders.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "hm_files.h"

namespace hm {
namespace pano {

struct SpatialTiff {
  float xpos;
  float ypos;
};

// Row-major host image of one-channel pixels.
template <typename T>
struct HostMat {
  int rows = 0, cols = 0;
  std::vector<T> data;
  HostMat() = default;
  HostMat(int r, int c, T fill = T{}) : rows(r), cols(c), data((size_t)r * c, fill) {}
  bool empty() const { return data.empty(); }
  T& at(int y, int x) { return data[(size_t)y * cols + x]; }
  const T& at(int y, int x) const { return data[(size_t)y * cols + x]; }
};

namespace detail {

// Loads `dir` under `kind`'s rules; fills maps, seam and placements.
inline bool load_control_dir(const std::string& dir, int kind, int n_images, std::vector<HostMat<uint16_t>>& cols,
                             std::vector<HostMat<uint16_t>>& rows, HostMat<uint8_t>& seam,
                             std::vector<SpatialTiff>& positions) {
  hmControlFiles* f = nullptr;
  if (hmControlFilesLoad(dir.c_str(), kind, n_images, &f) != HM_FILE_OK) return false;
  hmControlFilesInfo info;
  hmControlFilesGetInfo(f, &info);
  cols.assign(info.n_images, HostMat<uint16_t>());
  rows.assign(info.n_images, HostMat<uint16_t>());
  positions.clear();
  for (int i = 0; i < info.n_images; ++i) {
    const int w = (int)info.map_w[i], h = (int)info.map_h[i];
    cols[i] = HostMat<uint16_t>(h, w);
    rows[i] = HostMat<uint16_t>(h, w);
    memcpy(cols[i].data.data(), info.map_x[i], (size_t)w * h * sizeof(uint16_t));
    memcpy(rows[i].data.data(), info.map_y[i], (size_t)w * h * sizeof(uint16_t));
    positions.push_back(SpatialTiff{info.xpos[i], info.ypos[i]});
  }
  seam = HostMat<uint8_t>((int)info.seam_h, (int)info.seam_w);
  memcpy(seam.data.data(), info.seam, (size_t)info.seam_w * info.seam_h);
  hmControlFilesDestroy(f);
  return true;
}

}  // namespace detail

class ControlMasks {
 public:
  HostMat<uint16_t> img1_col, img1_row, img2_col, img2_row;
  HostMat<uint8_t> whole_seam_mask_image;
  std::vector<SpatialTiff> positions;

  ControlMasks() = default;
  // controlMasks.cpp:253-256: the caller checks is_valid().
  explicit ControlMasks(std::string game_dir) { (void)load(std::move(game_dir)); }
  // controlMasks.cpp:273-326; hmFilesLastError() says why a load failed.
  bool load(std::string game_dir) {
    std::vector<HostMat<uint16_t>> c, r;
    if (!detail::load_control_dir(game_dir, HM_CONTROL_TWO, 2, c, r, whole_seam_mask_image, positions)) return false;
    img1_col = std::move(c[0]);
    img1_row = std::move(r[0]);
    img2_col = std::move(c[1]);
    img2_row = std::move(r[1]);
    return true;
  }

  bool is_valid() const {
    return !img1_col.empty() && !img1_row.empty() && !img2_col.empty() && !img2_row.empty() &&
           !whole_seam_mask_image.empty() && positions.size() == 2;
  }
  // controlMasks.cpp:265-271
  size_t canvas_width() const {
    return (size_t)std::max(positions.at(0).xpos + img1_col.cols, positions.at(1).xpos + img2_col.cols);
  }
  size_t canvas_height() const {
    return (size_t)std::max(positions.at(0).ypos + img1_col.rows, positions.at(1).ypos + img2_col.rows);
  }
};

// ControlMasksN (src/pano/controlMasksN.h:13-38): N remap tables, the
// indexed seam with labels 0..N-1, N positions.
class ControlMasksN {
 public:
  std::vector<HostMat<uint16_t>> img_col, img_row;
  HostMat<uint8_t> whole_seam_mask_indexed;
  std::vector<SpatialTiff> positions;

  ControlMasksN() = default;
  ControlMasksN(const std::string& dir, int n_images) { (void)load(dir, n_images); }
  // controlMasksN.cpp:125-193
  bool load(const std::string& dir, int n_images) {
    return detail::load_control_dir(dir, HM_CONTROL_N, n_images, img_col, img_row, whole_seam_mask_indexed,
                                    positions);
  }

  bool is_valid() const {  // controlMasksN.cpp:195-203
    if (img_col.empty() || img_row.empty() || img_col.size() != img_row.size()) return false;
    if (positions.size() != img_col.size()) return false;
    return !whole_seam_mask_indexed.empty();
  }
  size_t canvas_width() const {  // controlMasksN.cpp:205-211
    float w = 0.0f;
    for (size_t i = 0; i < img_col.size(); ++i) w = std::max(w, positions[i].xpos + img_col[i].cols);
    return (size_t)w;
  }
  size_t canvas_height() const {  // controlMasksN.cpp:213-219
    float h = 0.0f;
    for (size_t i = 0; i < img_row.size(); ++i) h = std::max(h, positions[i].ypos + img_row[i].rows);
    return (size_t)h;
  }
};

// ControlMasks3 (src/pano/controlMasks3.h:12-43): three remap tables, the
// single-channel seam with labels 0..2, three positions.
class ControlMasks3 {
 public:
  HostMat<uint16_t> img0_col, img0_row, img1_col, img1_row, img2_col, img2_row;
  HostMat<uint8_t> whole_seam_mask_image;
  std::vector<SpatialTiff> positions;

  ControlMasks3() = default;
  explicit ControlMasks3(const std::string& game_dir) { (void)load(game_dir); }
  // controlMasks3.cpp:228-295
  bool load(const std::string& game_dir) {
    std::vector<HostMat<uint16_t>> c, r;
    if (!detail::load_control_dir(game_dir, HM_CONTROL_THREE, 3, c, r, whole_seam_mask_image, positions))
      return false;
    img0_col = std::move(c[0]);
    img0_row = std::move(r[0]);
    img1_col = std::move(c[1]);
    img1_row = std::move(r[1]);
    img2_col = std::move(c[2]);
    img2_row = std::move(r[2]);
    return true;
  }

  bool is_valid() const {  // controlMasks3.cpp:297-301
    return !img0_col.empty() && !img0_row.empty() && !img1_col.empty() && !img1_row.empty() && !img2_col.empty() &&
           !img2_row.empty() && !whole_seam_mask_image.empty() && positions.size() == 3;
  }
  size_t canvas_width() const {  // controlMasks3.cpp:303-309
    return (size_t)std::max({positions.at(0).xpos + img0_col.cols, positions.at(1).xpos + img1_col.cols,
                             positions.at(2).xpos + img2_col.cols});
  }
  size_t canvas_height() const {  // controlMasks3.cpp:311-317
    return (size_t)std::max({positions.at(0).ypos + img0_col.rows, positions.at(1).ypos + img1_col.rows,
                             positions.at(2).ypos + img2_col.rows});
  }
};

}  // namespace pano
}  // namespace hm
