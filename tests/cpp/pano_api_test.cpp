// C++ API test: drives hm::pano::cuda::CudaStitchPano, CudaStitchPanoN and
// CudaStitchPano3 through the reference's class interfaces (ControlMasks* ->
// ctor -> process(..., stream, canvas)), restating cases of
// src/pano/cudaPano_test.cpp, src/pano/cudaPano3_test.cpp and
// tests/python/test_pytorch_pano.py with uint8 pipelines, and the float tests
// of src/pano/cudaPano_test.cpp:141-204, cudaPanoN_test.cpp:192-299 and
// cudaPano3_test.cpp:133-196 with float pipelines.  Prints one line per case;
// exit code 0 = pass.  With a directory argument, the float canvases are also
// written there as raw float32 (<case>.f32, B x H x W x 4) for comparison with
// the oracle.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <memory>
#include <string>
#include <vector>

#include "cupano/pano/cudaPano.h"
#include "cupano/pano/cudaPano3.h"
#include "cupano/pano/cudaPanoN.h"

using hm::CudaMat;
using hm::pano::ControlMasks;
using hm::pano::ControlMasks3;
using hm::pano::ControlMasksN;
using hm::pano::cuda::CudaStitchPano3;
using hm::pano::cuda::CudaStitchPanoN;
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

void identity(int w, int h, HostMat<uint16_t>* mx, HostMat<uint16_t>* my) {
  *mx = HostMat<uint16_t>(h, w);
  *my = HostMat<uint16_t>(h, w);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      mx->at(y, x) = (uint16_t)x;
      my->at(y, x) = (uint16_t)y;
    }
}

// Three bands: label 0 in [0, w/3), 1 in [w/3, 2w/3), 2 beyond.
HostMat<uint8_t> bands3(int w, int h) {
  HostMat<uint8_t> seam(h, w, 0);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) seam.at(y, x) = (uint8_t)(x < w / 3 ? 0 : (x < 2 * w / 3 ? 1 : 2));
  return seam;
}

const uchar4 kColors[3] = {make_uchar4(200, 0, 0, 255), make_uchar4(0, 150, 0, 255), make_uchar4(0, 0, 100, 255)};

// Constant images per label: every level's blend selects the labelled colour,
// so away from the band edges the output is exactly that colour.
bool bands_ok(const std::vector<uchar4>& out, int w, int h, int margin) {
  if (out.size() != (size_t)w * h) return false;
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int lbl = x < w / 3 ? 0 : (x < 2 * w / 3 ? 1 : 2);
      const int e0 = lbl == 0 ? 0 : (lbl == 1 ? w / 3 : 2 * w / 3);
      const int e1 = lbl == 0 ? w / 3 : (lbl == 1 ? 2 * w / 3 : w);
      if (x - e0 < margin && e0 > 0) continue;
      if (e1 - x <= margin && e1 < w) continue;
      if (!same(out[(size_t)y * w + x], kColors[lbl])) return false;
    }
  return true;
}

std::vector<uchar4> download(CudaMat<uchar4>& m, int w, int h) {
  std::vector<uchar4> host((size_t)w * h);
  if (m.download(host.data()) != hipSuccess) return {};
  return host;
}

// ---- float pipelines ----
const char* g_dump_dir = nullptr;

void dump(const char* name, const std::vector<float4>& img) {
  if (!g_dump_dir) return;
  const std::string path = std::string(g_dump_dir) + "/" + name + ".f32";
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) return;
  std::fwrite(img.data(), sizeof(float4), img.size(), f);
  std::fclose(f);
}

// The reference tests' pattern as float, plus `frac` on the colour channels.
std::vector<float4> patterned_f(int w, int h, int idx, float frac) {
  const auto u = patterned(w, h, idx);
  std::vector<float4> img(u.size());
  for (size_t i = 0; i < u.size(); ++i) img[i] = make_float4(u[i].x + frac, u[i].y + frac, u[i].z + frac, u[i].w);
  return img;
}

float max_abs_diff(const std::vector<float4>& a, const std::vector<float4>& b) {
  if (a.empty() || a.size() != b.size()) return 1e30f;
  float d = 0.0f;
  for (size_t i = 0; i < a.size(); ++i) {
    d = std::max(d, std::fabs(a[i].x - b[i].x));
    d = std::max(d, std::fabs(a[i].y - b[i].y));
    d = std::max(d, std::fabs(a[i].z - b[i].z));
    d = std::max(d, std::fabs(a[i].w - b[i].w));
  }
  return d;
}

template <typename P>
std::vector<P> download_t(CudaMat<P>& m) {
  std::vector<P> host((size_t)m.width() * m.height());
  if (m.download(host.data()) != hipSuccess) return {};
  return host;
}

template <typename P>
std::vector<P> stitch2_f(const ControlMasks& masks, bool minimize, const std::vector<P>& a, const std::vector<P>& b,
                         int w, int h, int* level0_w) {
  CudaStitchPano<P, P> pano(1, 4, masks, /*quiet=*/true, minimize);
  if (!pano.status().ok()) return {};
  CudaMat<P> in1(a.data(), w, h), in2(b.data(), w, h);
  auto r = pano.process(in1, in2, nullptr,
                        std::make_unique<CudaMat<P>>(1, pano.canvas_width(), pano.canvas_height()));
  if (!r.ok()) return {};
  auto out = r.ConsumeValueOrDie();
  if (level0_w) {
    int levels = 0, lw[16] = {0}, lh[16] = {0};  // at most 16 levels
    *level0_w = hmStitchPanoLevels(pano.native_handle(), &levels, lw, lh) == 0 ? lw[0] : -1;
  }
  return download_t(*out);
}

std::vector<float4> stitchN_f(const ControlMasksN& masks, bool minimize, const std::vector<std::vector<float4>>& imgs,
                              int w, int h) {
  CudaStitchPanoN<float4, float4> pano(1, 4, masks, minimize, /*quiet=*/true);
  if (!pano.status().ok()) return {};
  std::vector<std::unique_ptr<CudaMat<float4>>> ins;
  std::vector<const CudaMat<float4>*> ptrs;
  for (const auto& im : imgs) {
    ins.push_back(std::make_unique<CudaMat<float4>>(im.data(), w, h));
    ptrs.push_back(ins.back().get());
  }
  auto r = pano.process(ptrs, nullptr,
                        std::make_unique<CudaMat<float4>>(1, pano.canvas_width(), pano.canvas_height()));
  if (!r.ok()) return {};
  auto out = r.ConsumeValueOrDie();
  return download_t(*out);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc > 1) g_dump_dir = argv[1];
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
  // src/pano/cudaPano_test.cpp:185-201: pyramid dumps of the full and
  // minimized blends; level 0 spans the canvas width only for the full one.
  {
    const int w = 384, h = 64, x2 = 192, cwid = w + x2;
    HostMat<uint8_t> seam(h, cwid, 0);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < cwid / 2; ++x) seam.at(y, x) = 1;
    const ControlMasks masks = make_masks(w, h, x2, seam);
    const auto i1 = patterned(w, h, 0), i2 = patterned(w, h, 1);
    CudaMat<uchar4> a(i1.data(), w, h), b(i2.data(), w, h);
    char tmpl[] = "/tmp/pano_dump_XXXXXX";
    const char* root = mkdtemp(tmpl);
    bool ok = root != nullptr;
    int lw[2] = {0, 0}, lh[2] = {0, 0};
    for (int mini = 0; ok && mini < 2; ++mini) {
      CudaStitchPano<uchar4, float4> pano(1, 4, masks, true, mini == 1);
      auto r = pano.process(a, b, nullptr,
                            std::make_unique<CudaMat<uchar4>>(1, pano.canvas_width(), pano.canvas_height()));
      const std::string dir = std::string(root) + (mini ? "/mini" : "/full");
      ok = r.ok() && pano.dump_soft_blend_pyramid(dir, nullptr).ok();
      FILE* f = ok ? std::fopen((dir + "/metadata.txt").c_str(), "r") : nullptr;
      char line[128];
      while (f && std::fgets(line, sizeof(line), f)) std::sscanf(line, "level_0=%dx%d", &lw[mini], &lh[mini]);
      if (f) std::fclose(f);
    }
    if (root) std::filesystem::remove_all(root);
    check(ok && lw[0] == cwid && lh[0] == h && lh[1] == h && lw[1] < lw[0],
          "dump_soft_blend_pyramid: full blend level 0 spans the canvas, the minimized one is narrower");
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
  // CudaStitchPanoN: three identity-mapped images, three label bands.
  {
    const int w = 96, h = 40;
    ControlMasksN masks;
    masks.img_col.resize(3);
    masks.img_row.resize(3);
    for (int i = 0; i < 3; ++i) identity(w, h, &masks.img_col[i], &masks.img_row[i]);
    masks.whole_seam_mask_indexed = bands3(w, h);
    masks.positions = {SpatialTiff{0, 0}, SpatialTiff{0, 0}, SpatialTiff{0, 0}};
    std::vector<std::vector<uchar4>> imgs;
    for (int i = 0; i < 3; ++i) imgs.push_back(constant(w, h, kColors[i]));
    std::vector<std::unique_ptr<CudaMat<uchar4>>> ins;
    for (int i = 0; i < 3; ++i) ins.push_back(std::make_unique<CudaMat<uchar4>>(imgs[i].data(), w, h));
    std::vector<const CudaMat<uchar4>*> in_ptrs = {ins[0].get(), ins[1].get(), ins[2].get()};
    for (int levels : {0, 4}) {
      CudaStitchPanoN<uchar4, float4> pano(1, levels, masks, /*minimize_blend=*/true, /*quiet=*/true);
      bool ok = pano.status().ok() && pano.canvas_width() == w && pano.canvas_height() == h;
      if (ok) {
        auto r = pano.process(in_ptrs, nullptr, std::make_unique<CudaMat<uchar4>>(1, w, h));
        ok = r.ok();
        if (ok) {
          auto out = r.ConsumeValueOrDie();
          ok = bands_ok(download(*out, w, h), w, h, levels == 0 ? 0 : 16);
        }
      }
      check(ok, levels == 0 ? "CudaStitchPanoN hard seam selects each labelled image"
                            : "CudaStitchPanoN soft seam of constant images keeps each colour away from the edges");
    }
    CudaStitchPanoN<uchar4, float4> pano(1, 2, masks, true, true);
    std::vector<const CudaMat<uchar4>*> two = {ins[0].get(), ins[1].get()};
    auto bad = pano.process(two, nullptr, std::make_unique<CudaMat<uchar4>>(1, w, h));
    check(!bad.ok() && pano.status().ok(), "CudaStitchPanoN input-count errors are not sticky");
  }
  // CudaStitchPano3: src/pano/cudaPano3_test.cpp:133-196 (1x1, one-hot label 1 -> image 1), both paths.
  {
    ControlMasks3 masks;
    for (auto* m : {&masks.img0_col, &masks.img0_row, &masks.img1_col, &masks.img1_row, &masks.img2_col,
                    &masks.img2_row})
      *m = HostMat<uint16_t>(1, 1, 0);
    masks.whole_seam_mask_image = HostMat<uint8_t>(1, 1, 1);
    masks.positions = {SpatialTiff{0, 0}, SpatialTiff{0, 0}, SpatialTiff{0, 0}};
    const uchar4 v[3] = {make_uchar4(30, 30, 30, 255), make_uchar4(60, 60, 60, 255), make_uchar4(90, 90, 90, 255)};
    CudaMat<uchar4> i0(&v[0], 1, 1), i1(&v[1], 1, 1), i2(&v[2], 1, 1);
    bool ok = true;
    for (bool fused : {true, false}) {
      CudaStitchPano3<uchar4, float4> pano(1, 1, masks, true);
      auto r = pano.process(i0, i1, i2, nullptr, std::make_unique<CudaMat<uchar4>>(1, 1, 1), fused);
      if (!r.ok()) {
        ok = false;
        continue;
      }
      auto out = r.ConsumeValueOrDie();
      const auto host = download(*out, 1, 1);
      ok = ok && host.size() == 1 && same(host[0], v[1]);
    }
    check(ok, "CudaStitchPano3 one-hot label 1 selects image 1 (fused and unfused)");
  }
  // src/pano/cudaPano_test.cpp:141-204 with a float pipeline: the minimized and
  // full-canvas blends agree within 1e-4 and the minimized level 0 is narrower.
  {
    const int w = 384, h = 64, x2 = 192, cwid = w + x2;
    HostMat<uint8_t> seam(h, cwid, 0);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < cwid / 2; ++x) seam.at(y, x) = 1;
    const ControlMasks masks = make_masks(w, h, x2, seam);
    for (float frac : {0.0f, 0.375f}) {
      const auto i1 = patterned_f(w, h, 0, frac), i2 = patterned_f(w, h, 1, frac);
      int lw_full = 0, lw_mini = 0;
      const auto full = stitch2_f(masks, false, i1, i2, w, h, &lw_full);
      const auto mini = stitch2_f(masks, true, i1, i2, w, h, &lw_mini);
      dump(frac == 0.0f ? "pano2_float_mini" : "pano2_float_mini_frac", mini);
      check(full.size() == (size_t)cwid * h && max_abs_diff(full, mini) <= 1e-4f && lw_full == cwid &&
                lw_mini > 0 && lw_mini < cwid,
            frac == 0.0f ? "CudaStitchPano<float4, float4>: minimized equals full within 1e-4, narrower level 0"
                         : "CudaStitchPano<float4, float4> with fractional inputs: minimized equals full within 1e-4");
    }
    // float3: the same blend on three channels equals the float4 one's colour channels.
    std::vector<float3> a3, b3;
    const auto i1 = patterned_f(w, h, 0, 0.375f), i2 = patterned_f(w, h, 1, 0.375f);
    for (size_t i = 0; i < i1.size(); ++i) {
      a3.push_back(make_float3(i1[i].x, i1[i].y, i1[i].z));
      b3.push_back(make_float3(i2[i].x, i2[i].y, i2[i].z));
    }
    const auto full3 = stitch2_f(masks, false, a3, b3, w, h, nullptr);
    const auto mini3 = stitch2_f(masks, true, a3, b3, w, h, nullptr);
    std::vector<float4> full4(full3.size()), mini4(mini3.size());
    for (size_t i = 0; i < full3.size(); ++i) full4[i] = make_float4(full3[i].x, full3[i].y, full3[i].z, 0.0f);
    for (size_t i = 0; i < mini3.size(); ++i) mini4[i] = make_float4(mini3[i].x, mini3[i].y, mini3[i].z, 0.0f);
    dump("pano2_float3_mini_frac", mini4);
    check(full3.size() == (size_t)cwid * h && max_abs_diff(full4, mini4) <= 1e-4f,
          "CudaStitchPano<float3, float3>: minimized equals full within 1e-4");
  }
  // src/pano/cudaPanoN_test.cpp:192-299 with a float pipeline: two-way overlap,
  // no overlap, a gap (transparent uncovered columns) and a four-way seam
  // intersection; minimized and full blends agree within 1e-4.
  {
    struct NCase {
      const char* name;
      int w, h, n, cw;
      int xs[4];
      int band0, band1;  // columns of label 1; band0 < 0: four quadrant labels
    };
    const NCase cases[] = {
        {"panoN_float_two_way_overlap", 384, 384, 2, 384, {0, 0, 0, 0}, 192, 384},
        {"panoN_float_no_overlap", 192, 384, 2, 384, {0, 192, 0, 0}, 192, 384},
        {"panoN_float_gap", 128, 384, 2, 384, {0, 256, 0, 0}, 192, 384},
        {"panoN_float_multi_seam_intersection", 384, 384, 4, 384, {0, 0, 0, 0}, -1, -1},
    };
    for (const NCase& c : cases) {
      ControlMasksN masks;
      masks.img_col.resize(c.n);
      masks.img_row.resize(c.n);
      std::vector<std::vector<float4>> imgs;
      for (int i = 0; i < c.n; ++i) {
        identity(c.w, c.h, &masks.img_col[i], &masks.img_row[i]);
        masks.positions.push_back(SpatialTiff{(float)c.xs[i], 0.0f});
        imgs.push_back(patterned_f(c.w, c.h, i, 0.0f));
      }
      HostMat<uint8_t> seam(c.h, c.cw, 0);
      if (c.band0 >= 0) {
        for (int y = 0; y < c.h; ++y)
          for (int x = c.band0; x < c.band1; ++x) seam.at(y, x) = 1;
      } else {
        const int cx = c.w / 2, cy = c.h / 2;
        for (int y = cy - 32; y < cy + 32; ++y)
          for (int x = cx - 32; x < cx + 32; ++x) seam.at(y, x) = (uint8_t)((y >= cy ? 2 : 0) + (x >= cx ? 1 : 0));
      }
      masks.whole_seam_mask_indexed = seam;
      const auto full = stitchN_f(masks, false, imgs, c.w, c.h);
      const auto mini = stitchN_f(masks, true, imgs, c.w, c.h);
      dump(c.name, mini);
      bool ok = full.size() == (size_t)c.cw * c.h && max_abs_diff(full, mini) <= 1e-4f;
      if (ok && c.xs[1] == 256)
        for (int y = 0; y < c.h; ++y)
          for (int x = 128; x < 256; ++x) ok = ok && mini[(size_t)y * c.cw + x].w == 0.0f;
      check(ok, c.name);
    }
  }
  // src/pano/cudaPano3_test.cpp:133-196 with a float pipeline: a 1x1 one-hot
  // label 1 selects image 1's value (within 1e-3), fused and unfused.
  {
    ControlMasks3 masks;
    for (auto* m : {&masks.img0_col, &masks.img0_row, &masks.img1_col, &masks.img1_row, &masks.img2_col,
                    &masks.img2_row})
      *m = HostMat<uint16_t>(1, 1, 0);
    masks.whole_seam_mask_image = HostMat<uint8_t>(1, 1, 1);
    masks.positions = {SpatialTiff{0, 0}, SpatialTiff{0, 0}, SpatialTiff{0, 0}};
    const float4 v[3] = {make_float4(30, 30, 30, 255), make_float4(60, 60, 60, 255), make_float4(90, 90, 90, 255)};
    CudaMat<float4> i0(&v[0], 1, 1), i1(&v[1], 1, 1), i2(&v[2], 1, 1);
    bool ok = true;
    for (bool fused : {true, false}) {
      CudaStitchPano3<float4, float4> pano(1, 1, masks, true);
      auto r = pano.process(i0, i1, i2, nullptr, std::make_unique<CudaMat<float4>>(1, 1, 1), fused);
      if (!r.ok()) {
        ok = false;
        continue;
      }
      auto out = r.ConsumeValueOrDie();
      const auto host = download_t(*out);
      ok = ok && host.size() == 1 && max_abs_diff(host, {v[1]}) <= 1e-3f;
    }
    check(ok, "CudaStitchPano3<float4, float4> one-hot label 1 selects image 1 (fused and unfused)");
  }
  std::printf("%d failure(s)\n", g_failures);
  return g_failures ? 1 : 0;
}
