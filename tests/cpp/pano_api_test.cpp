// C++ API test: drives hm::pano::cuda::CudaStitchPano through the reference's
// class interface (ControlMasks -> ctor -> process(in1, in2, stream, canvas)),
// restating src/pano/cudaPano_test.cpp and tests/python/test_pytorch_pano.py
// cases with a uint8 pipeline.  Prints one line per case; exit code 0 = pass.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "cupano/pano/cudaPano.h"

using hm::CudaMat;
using hm::pano::ControlMasks;
using hm::pano::HostMat;
using hm::pano::SpatialTiff;
using hm::pano::cuda::CudaStitchPano;

namespace {

int g_failures = 0;

void check(bool ok, const char* what) {
  std::printf("%s: %s\n", ok ? "PASS" : "FAIL", what);
  if (!ok) ++g_failures;
}

ControlMasks make_masks(int w, int h, int x2, const HostMat<uint8_t>& seam) {
  ControlMasks m;
  for (auto* mx : {&m.img1_col, &m.img2_col}) {
    *mx = HostMat<uint16_t>(h, w);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) mx->at(y, x) = (uint16_t)x;
  }
  for (auto* my : {&m.img1_row, &m.img2_row}) {
    *my = HostMat<uint16_t>(h, w);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) my->at(y, x) = (uint16_t)y;
  }
  m.whole_seam_mask_image = seam;
  m.positions = {SpatialTiff{0.0f, 0.0f}, SpatialTiff{(float)x2, 0.0f}};
  return m;
}

std::vector<uchar4> constant(int w, int h, uchar4 v) { return std::vector<uchar4>((size_t)w * h, v); }

std::vector<uchar4> patterned(int w, int h, int idx) {
  std::vector<uchar4> img((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int base = idx * 50;
      img[(size_t)y * w + x] = make_uchar4(base + x % 17, base + y % 19, base + (x + y) % 23, 255);
    }
  return img;
}

std::vector<uchar4> stitch(const ControlMasks& masks, int levels, bool minimize, const std::vector<uchar4>& a,
                           const std::vector<uchar4>& b, int w, int h, int* cw, int* ch) {
  CudaStitchPano<uchar4, float4> pano(1, levels, masks, /*quiet=*/true, minimize);
  if (!pano.status().ok()) {
    std::printf("ctor failed: %s\n", pano.status().message().c_str());
    return {};
  }
  *cw = pano.canvas_width();
  *ch = pano.canvas_height();
  CudaMat<uchar4> in1(a.data(), w, h), in2(b.data(), w, h);
  auto canvas = std::make_unique<CudaMat<uchar4>>(1, *cw, *ch);
  hipStream_t stream = nullptr;
  auto result = pano.process(in1, in2, stream, std::move(canvas));
  if (!result.ok()) {
    std::printf("process failed: %s\n", result.status().message().c_str());
    return {};
  }
  std::unique_ptr<CudaMat<uchar4>> out = result.ConsumeValueOrDie();
  std::vector<uchar4> host((size_t)(*cw) * (*ch));
  if (out->download(host.data()) != hipSuccess) return {};
  return host;
}

bool same(uchar4 a, uchar4 b) { return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w; }

}  // namespace

int main() {
  // tests/python/test_pytorch_pano.py:109-152: hard seam and single-level soft seam select exactly.
  {
    const int w = 256, h = 32, x2 = 128, cwid = w + x2;
    HostMat<uint8_t> seam(h, cwid, 0);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < 192; ++x) seam.at(y, x) = 1;
    const ControlMasks masks = make_masks(w, h, x2, seam);
    const auto i1 = constant(w, h, make_uchar4(10, 20, 30, 255));
    const auto i2 = constant(w, h, make_uchar4(100, 110, 120, 255));
    for (int levels : {0, 1}) {
      int cw = 0, ch = 0;
      const auto out = stitch(masks, levels, true, i1, i2, w, h, &cw, &ch);
      bool ok = !out.empty() && cw == cwid && ch == h;
      for (int y = 0; ok && y < h; ++y)
        for (int x = 0; ok && x < cwid; ++x) {
          const uchar4 want = x < 192 ? i1[(size_t)y * w + x] : i2[(size_t)y * w + (x - x2)];
          ok = same(out[(size_t)y * cwid + x], want);
        }
      check(ok, levels == 0 ? "hard seam selects image 1 left of the seam and image 2 right of it"
                            : "single-level soft seam equals the binary selection");
    }
  }
  // src/pano/cudaPano_test.cpp:141-204: minimized and full-canvas blends agree (1 LSB after truncation).
  {
    const int w = 384, h = 64, x2 = 192, cwid = w + x2;
    HostMat<uint8_t> seam(h, cwid, 0);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < cwid / 2; ++x) seam.at(y, x) = 1;
    const ControlMasks masks = make_masks(w, h, x2, seam);
    const auto i1 = patterned(w, h, 0), i2 = patterned(w, h, 1);
    int cw = 0, ch = 0;
    const auto full = stitch(masks, 4, false, i1, i2, w, h, &cw, &ch);
    const auto mini = stitch(masks, 4, true, i1, i2, w, h, &cw, &ch);
    int maxd = full.empty() || mini.empty() ? 999 : 0;
    for (size_t i = 0; maxd < 999 && i < full.size(); ++i) {
      maxd = std::max(maxd, std::abs((int)full[i].x - (int)mini[i].x));
      maxd = std::max(maxd, std::abs((int)full[i].y - (int)mini[i].y));
      maxd = std::max(maxd, std::abs((int)full[i].z - (int)mini[i].z));
      maxd = std::max(maxd, std::abs((int)full[i].w - (int)mini[i].w));
    }
    check(maxd <= 1, "minimize_blend matches the full-canvas blend within 1 LSB");
  }
  // Constructor errors are sticky (cudaPano.inl:28-31, 460-472).
  {
    ControlMasks empty;
    CudaStitchPano<uchar4, float4> pano(1, 6, empty, true);
    check(!pano.status().ok() && pano.status().code() == hipErrorFileNotFound,
          "invalid control masks give a sticky FileNotFound status");
    CudaMat<uchar4> a(1, 4, 4), b(1, 4, 4);
    auto r = pano.process(a, b, nullptr, std::make_unique<CudaMat<uchar4>>(1, 4, 4));
    check(!r.ok() && r.status().code() == hipErrorFileNotFound, "process returns the constructor's status");
  }
  std::printf("%d failure(s)\n", g_failures);
  return g_failures ? 1 : 0;
}
