// Loads a rig control directory through the C++ ControlMasks classes and
// prints a digest (sizes, placements, canvas size, checksums) that the
// Python test compares with the oracle.  Host code only; runs without a GPU.
//
//   control_files_test <dir> two|three|n <n_images>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "cupano/pano/controlMasks.h"

using hm::pano::HostMat;

template <typename T>
static unsigned long long checksum(const HostMat<T>& m) {
  unsigned long long h = 1469598103934665603ull;  // FNV-1a over the values
  for (T v : m.data) {
    h ^= (unsigned long long)v;
    h *= 1099511628211ull;
  }
  return h;
}

static void print_map(int i, const HostMat<uint16_t>& x, const HostMat<uint16_t>& y) {
  printf("map %d %d %d %llu %llu\n", i, x.cols, x.rows, checksum(x), checksum(y));
}

static void print_positions(const std::vector<hm::pano::SpatialTiff>& p) {
  for (size_t i = 0; i < p.size(); ++i) printf("pos %zu %.9g %.9g\n", i, p[i].xpos, p[i].ypos);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <dir> two|three|n [n_images]\n", argv[0]);
    return 2;
  }
  const std::string dir = argv[1], kind = argv[2];
  if (kind == "two") {
    hm::pano::ControlMasks m(dir);
    if (!m.is_valid()) {
      printf("error %s\n", hmFilesLastError());
      return 1;
    }
    print_map(0, m.img1_col, m.img1_row);
    print_map(1, m.img2_col, m.img2_row);
    print_positions(m.positions);
    printf("seam %d %d %llu\n", m.whole_seam_mask_image.cols, m.whole_seam_mask_image.rows,
           checksum(m.whole_seam_mask_image));
    printf("canvas %zu %zu\n", m.canvas_width(), m.canvas_height());
  } else if (kind == "three") {
    hm::pano::ControlMasks3 m(dir);
    if (!m.is_valid()) {
      printf("error %s\n", hmFilesLastError());
      return 1;
    }
    print_map(0, m.img0_col, m.img0_row);
    print_map(1, m.img1_col, m.img1_row);
    print_map(2, m.img2_col, m.img2_row);
    print_positions(m.positions);
    printf("seam %d %d %llu\n", m.whole_seam_mask_image.cols, m.whole_seam_mask_image.rows,
           checksum(m.whole_seam_mask_image));
    printf("canvas %zu %zu\n", m.canvas_width(), m.canvas_height());
  } else {
    const int n = argc > 3 ? atoi(argv[3]) : 0;
    hm::pano::ControlMasksN m;
    if (!m.load(dir, n) || !m.is_valid()) {
      printf("error %s\n", hmFilesLastError());
      return 1;
    }
    for (int i = 0; i < n; ++i) print_map(i, m.img_col[i], m.img_row[i]);
    print_positions(m.positions);
    printf("seam %d %d %llu\n", m.whole_seam_mask_indexed.cols, m.whole_seam_mask_indexed.rows,
           checksum(m.whole_seam_mask_indexed));
    printf("canvas %zu %zu\n", m.canvas_width(), m.canvas_height());
  }
  return 0;
}
