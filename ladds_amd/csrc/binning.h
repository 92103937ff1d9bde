// binning.h -- cell assignment shared by the propagation (count pass) and the scatter.
//
// Cells have side s >= 2 * R, R the search radius (the cutoff in the FP64 form, the
// inflated FP32 prefilter radius in the split form).  An object at fractional cell
// position f < 1/2 (per axis) can only have partners within R in its own cell or the
// one below, otherwise in its own cell or the one above; its search block is the
// 2x2x2 cells starting at (ix - [f < 1/2], ...).  Any pair within R lies in each
// other's search block, so scanning the 8 buckets of the block from the lower sorted
// position finds every pair exactly once.  The split form bins the FP32-rounded
// position it also stores in its record, so binning and search use the same point.
#pragma once

#include <hip/hip_runtime.h>

#include "engine.h"

namespace ladds {

// autopas::utils::ArrayMath::dot: accumulate from 0, no contraction: ((0 + a0 b0) + a1 b1) + a2 b2.
__device__ __forceinline__ double exact_dot3(double a0, double a1, double a2, double b0, double b1, double b2) {
#pragma clang fp contract(off)
  double s = 0.0;
  s += a0 * b0;
  s += a1 * b1;
  s += a2 * b2;
  return s;
}

struct CellInfo {
  uint32_t bucket;  // folded bucket of the object's own cell
  uint32_t base;    // folded lower corner of its 2x2x2 search block, packed (fx, fy, fz)
};

// The point an object is binned by: its FP64 position, or its FP32-rounded position
// in the split form.
template <bool F32>
__device__ __forceinline__ void binning_point(double &x, double &y, double &z) {
  if constexpr (F32) {
    x = (double)(float)x;
    y = (double)(float)y;
    z = (double)(float)z;
  }
}

__device__ __forceinline__ int32_t cell_axis(double x, double inv_cell, uint32_t &low) {
  double u = x * inv_cell;
  u = fmin(fmax(u, -1.0e9), 1.0e9);
  const double fl = floor(u);
  low = (u - fl < 0.5) ? 1u : 0u;
  return (int32_t)fl;
}

__device__ __forceinline__ CellInfo cell_of(double x, double y, double z, const GridParams &g) {
  uint32_t lx, ly, lz;
  const int32_t ix = cell_axis(x, g.inv_cell_x, lx);
  const int32_t iy = cell_axis(y, g.inv_cell, ly);
  const int32_t iz = cell_axis(z, g.inv_cell, lz);
  const uint32_t mx = (1u << g.bx) - 1u, my = (1u << g.by) - 1u, mz = (1u << g.bz) - 1u;
  const uint32_t sy = g.bz, sx = g.by + g.bz;
  CellInfo c;
  c.bucket = (((uint32_t)ix & mx) << sx) | (((uint32_t)iy & my) << sy) | ((uint32_t)iz & mz);
  c.base = (((uint32_t)(ix - (int32_t)lx) & mx) << sx) | (((uint32_t)(iy - (int32_t)ly) & my) << sy) |
           ((uint32_t)(iz - (int32_t)lz) & mz);
  return c;
}

// Folded in-plane bucket of the plane tiles from y and z only: rows of height 1 / inv_cell_y
// along y and buckets of width 1 / inv_cell_z along z (each the cell size, or twice it).
// low_y (low_z) = 1 when the point lies in the lower half of its row (bucket) along y (z).
__device__ __forceinline__ uint32_t yz_bucket(double y, double z, const GridParams &g, uint32_t &low_y,
                                              uint32_t &low_z) {
  const int32_t iy = cell_axis(y, g.inv_cell_y, low_y);
  const int32_t iz = cell_axis(z, g.inv_cell_z, low_z);
  const uint32_t my = (1u << g.by) - 1u, mz = (1u << g.bz) - 1u;
  return (((uint32_t)iy & my) << g.bz) | ((uint32_t)iz & mz);
}

__device__ __forceinline__ uint32_t yz_bucket(double y, double z, const GridParams &g) {
  uint32_t ly, lz;
  return yz_bucket(y, z, g, ly, lz);
}

}  // namespace ladds
