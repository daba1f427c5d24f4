// controlMasks.h -- in-memory control data of a rig (no file I/O).
//
// The fields of hm-cupano's ControlMasks (src/pano/controlMasks.h:30-66):
// img1_col/img1_row/img2_col/img2_row are the CV_16U remap tables,
// whole_seam_mask_image is the seam AFTER load_seam_mask's inversion
// (controlMasks.cpp:225-251: 1 selects image 1), positions are the normalised
// SpatialTiff placements.  cv::Mat is replaced by a minimal host image.
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

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

class ControlMasks {
 public:
  HostMat<uint16_t> img1_col, img1_row, img2_col, img2_row;
  HostMat<uint8_t> whole_seam_mask_image;
  std::vector<SpatialTiff> positions;

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

}  // namespace pano
}  // namespace hm
