// hm_pano2_device.hpp -- the 2-image class's remap kernels (level-0 blend
// inputs with the canvas outside the blend, and the hard seam), as bodies
// shared by their own launches and the batched pass.
#pragma once

#include "hm_pano_device.hpp"

namespace hmdev {

struct Pano2L0 {
  Surf src[2];
  const uint32_t* map[2];
  int mw[2], mh[2], px[2], py[2];
  uint8_t* canvas;
  uint32_t cpitch;
  int cw, ch;
  int write_canvas, skip_x0, skip_x1;  // canvas columns [skip_x0, skip_x1) are left to the blend write-back
  uint8_t* g0[2];
  int bw, bh;
  int d[2], ulo[2], uhi[2];  // G_i[u] <- canvas column u + d[i] for u in [ulo[i], uhi[i])
  int g2_canvas;             // 1: G2 takes the canvas after both remaps; 0: image 2 alone
  // Canvas column spans of one launch, [span_x0[k], span_x1[k]), cut into
  // span_tiles[k] tiles of kBlock.x * PX columns.
  int span_x0[2], span_x1[2], span_tiles[2];
  int tiles_x, tiles_y;  // tiles_x = span_tiles[0] + span_tiles[1]
  int xcd;               // 1: tiles in horizontal bands, one per XCD (its L2 keeps the band's source lines)
};

// The canvas after both remaps (cudaPano.inl:210-303) outside the blend
// columns, and the two level-0 blend inputs: G1 = canvas after remap 1 and
// G2 = canvas after remap 2, which overwrites image 2's whole box (the
// make_full copies, cudaPano.inl:225-236, 306-317), or image 2 alone when the
// blend region is the whole canvas.
template <int C, int PX, int RY>
__device__ __forceinline__ void level0_tile(const Pano2L0& a, int t, int b) {
  const int txa = t % a.tiles_x;
  const int k = txa < a.span_tiles[0] ? 0 : 1;
  const int tx = txa - (k ? a.span_tiles[0] : 0);
  const int X0 = a.span_x0[k] + tx * (int)blockDim.x * PX + (int)threadIdx.x;
  const int Xend = a.span_x1[k];
  // RY rows per thread, blockDim.y apart: a tile is blockDim.x * PX wide and
  // blockDim.y * RY tall, so the source lines that neighbouring canvas rows
  // share are read while the tile is resident.
  const int Y0 = (t / a.tiles_x) * (int)blockDim.y * RY + (int)threadIdx.y;
  uint32_t m[RY][2][PX];
  bool valid[RY][2][PX];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int Y = Y0 + r * (int)blockDim.y;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ry = Y - a.py[i];
      const bool row_ok = Y < a.ch && ry >= 0 && ry < a.mh[i];
      const uint32_t* mrow = a.map[i] + (size_t)(row_ok ? ry : 0) * a.mw[i];
#pragma unroll
      for (int j = 0; j < PX; ++j) {
        const int rx = X0 + j * (int)blockDim.x - a.px[i];
        valid[r][i][j] = row_ok && rx >= 0 && rx < a.mw[i] && X0 + j * (int)blockDim.x < Xend;
        m[r][i][j] = valid[r][i][j] ? mrow[rx] : 0u;
      }
    }
  }
  uint32_t v[RY][2][PX];
#pragma unroll
  for (int r = 0; r < RY; ++r)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < PX; ++j) v[r][i][j] = remap_px_sel<C>(a.src[i], b, m[r][i][j], valid[r][i][j]);
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int Y = Y0 + r * (int)blockDim.y;
    if (Y >= a.ch) break;
    uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
    uint8_t* g0row = a.g0[0] + ((size_t)b * a.bh + Y) * a.bw * C;
    uint8_t* g1row = a.g0[1] + ((size_t)b * a.bh + Y) * a.bw * C;
#pragma unroll
    for (int j = 0; j < PX; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      if (X >= Xend) continue;
      const uint32_t c1 = valid[r][0][j] ? v[r][0][j] : 0u;
      const uint32_t c12 = valid[r][1][j] ? v[r][1][j] : c1;
      if (a.write_canvas && !(X >= a.skip_x0 && X < a.skip_x1)) store_px<C>(crow + (size_t)X * C, c12);
      if (Y < a.bh) {
        const int u1 = X - a.d[0];
        if (u1 >= a.ulo[0] && u1 < a.uhi[0]) store_px<C>(g0row + (size_t)u1 * C, c1);
        const int u2 = X - a.d[1];
        if (u2 >= a.ulo[1] && u2 < a.uhi[1])
          store_px<C>(g1row + (size_t)u2 * C, a.g2_canvas ? c12 : (valid[r][1][j] ? v[r][1][j] : 0u));
      }
    }
  }
}

// One tile per workgroup, or (a.grid_tiles set) a persistent grid that walks
// the tiles with a stride: a smaller grid holds fewer CU slots and memory
// requests, leaving room to latency-bound work on another stream.
// grid_x: the workgroups of the launch (the tile stride).  Runs on 64 x 4
// threads (kBlock).
template <int C, int PX, int RY>
__device__ __forceinline__ void pano2_level0_body(const Pano2L0& a, Blk blk, int grid_x) {
  const int total = a.tiles_x * a.tiles_y;
  for (int tt = blk.x; tt < total; tt += grid_x) level0_tile<C, PX, RY>(a, a.xcd ? xcd_tile(tt, total) : tt, blk.z);
}

template <int C, int PX, int RY>
__global__ void __launch_bounds__(256) pano2_level0_kernel(Pano2L0 a) {
  (void)a;
  pano2_level0_body<C, PX, RY>(kernel_arg<Pano2L0>(), hw_blk(), (int)gridDim.x);
}


struct Pano2Hard {
  Surf src[2];
  const uint32_t* map[2];
  int mw[2], mh[2], px[2], py[2];
  uint8_t* canvas;
  uint32_t cpitch;
  int cw, ch;
  const uint8_t* seam;  // canvas-sized, 1 selects image 1, 0 image 2
  int tiles_x, tiles_y;
};

// Hard seam: the two dest-map remaps of cudaPano.inl:262-276 and :345-359
// (image 1 where seam == 1, image 2 where seam == 0) plus the canvas memset,
// as one pass; PX pixels per thread as above.
// Runs on 64 x 4 threads (kBlock).
template <int C, int PX, int RY>
__device__ __forceinline__ void pano2_hard_body(const Pano2Hard& a, Blk blk) {
  const int t = xcd_tile(blk.x, a.tiles_x * a.tiles_y);
  const int X0 = (t % a.tiles_x) * (int)blockDim.x * PX + (int)threadIdx.x;
  const int Y0 = (t / a.tiles_x) * (int)blockDim.y * RY + (int)threadIdx.y;
  const int b = blk.z;
  // Three rounds of loads, each issued in full before the next: the seam
  // bytes, the selected maps, the source pixels.  Loads outside the canvas or
  // a map go to element 0 (a valid address) and are discarded.
  uint32_t sv[RY][PX];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int Y = Y0 + r * (int)blockDim.y;
    const uint8_t* srow = a.seam + (size_t)(Y < a.ch ? Y : 0) * a.cw;
#pragma unroll
    for (int j = 0; j < PX; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      sv[r][j] = srow[X < a.cw ? X : 0];
    }
  }
  // Both images' geometry pinned in scalar registers: the per-pixel choice
  // between them is a select, not a per-lane load of an argument field.
  const int px0 = pin_sgpr(a.px[0]), px1 = pin_sgpr(a.px[1]);
  const int py0 = pin_sgpr(a.py[0]), py1 = pin_sgpr(a.py[1]);
  const int mw0 = pin_sgpr(a.mw[0]), mw1 = pin_sgpr(a.mw[1]);
  const int mh0 = pin_sgpr(a.mh[0]), mh1 = pin_sgpr(a.mh[1]);
  const uint32_t* map0 = pin_sgpr(a.map[0]);
  const uint32_t* map1 = pin_sgpr(a.map[1]);
  const Surf src0 = pin_sgpr(a.src[0]), src1 = pin_sgpr(a.src[1]);
  uint32_t m[RY][PX];
  int sel[RY][PX];  // 0: image 1, 1: image 2, -1: nothing
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int Y = Y0 + r * (int)blockDim.y;
#pragma unroll
    for (int j = 0; j < PX; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      const bool first = sv[r][j] == 1;
      const int rx = X - (first ? px0 : px1);
      const int ry = Y - (first ? py0 : py1);
      const int mw = first ? mw0 : mw1;
      const int mh = first ? mh0 : mh1;
      const bool hit = X < a.cw && Y < a.ch && sv[r][j] <= 1 && rx >= 0 && ry >= 0 && rx < mw && ry < mh;
      sel[r][j] = hit ? (first ? 0 : 1) : -1;
      m[r][j] = (first ? map0 : map1)[hit ? (size_t)ry * mw + rx : 0];
    }
  }
  uint32_t v[RY][PX];
#pragma unroll
  for (int r = 0; r < RY; ++r)
#pragma unroll
    for (int j = 0; j < PX; ++j)
      v[r][j] = remap_px_dm_sel<C>(sel[r][j] == 0 ? src0 : src1, b, m[r][j], sel[r][j] >= 0);
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int Y = Y0 + r * (int)blockDim.y;
    if (Y >= a.ch) break;
    uint8_t* crow = a.canvas + ((size_t)b * a.ch + Y) * a.cpitch;
#pragma unroll
    for (int j = 0; j < PX; ++j) {
      const int X = X0 + j * (int)blockDim.x;
      if (X < a.cw) store_px<C>(crow + (size_t)X * C, v[r][j]);
    }
  }
}

template <int C, int PX, int RY>
__global__ void __launch_bounds__(256) pano2_hard_kernel(Pano2Hard a) {
  (void)a;
  pano2_hard_body<C, PX, RY>(kernel_arg<Pano2Hard>(), hw_blk());
}

}  // namespace hmdev
