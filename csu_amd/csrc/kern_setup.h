// kern_setup.h -- set-up kernels run once per mesh.
#pragma once
#include "kern_common.h"

namespace csu {

// HBM streaming reference: y = s x over n doubles, two per lane (16-byte
// accesses), one pair per lane over a grid that covers the arrays, with
// non-temporal loads and stores.  Of the forms measured on MI355X (grid-
// stride loops of 8-64 workgroups per CU: 4.4-4.9 TB/s read + write; one
// pair per lane: 6.2 TB/s; four pairs per lane: 5.8 TB/s; a device-to-device
// hipMemcpy: 5.3 TB/s) this one is the fastest, 6.7 TB/s: the practical
// read + write ceiling the stage kernels' bytes are compared with.
__global__ __launch_bounds__(CSU_BLOCK) void k_stream(const double2* __restrict__ x, double2* __restrict__ y,
                                                      size_t n2, double s) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) {
    const double a = __builtin_nontemporal_load(&x[i].x), b = __builtin_nontemporal_load(&x[i].y);
    __builtin_nontemporal_store(a * s, &y[i].x);
    __builtin_nontemporal_store(b * s, &y[i].y);
  }
}

// FP32 copy of an FP64 array (round to nearest, as a host conversion does).
__global__ __launch_bounds__(CSU_BLOCK) void k_to_f32(const double* __restrict__ src, float* __restrict__ dst,
                                                      size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// saCalcD (sa.c:703-763): distance from each cell's vertex-average centre to
// the nearest wall / wallT segment (point-to-segment, parameter clamped to
// [0, 1]).  Segments (p0x, p0y, p1x, p1y) are SoA [4][nw] and are staged
// through LDS one block-sized tile at a time; every cell of the block scans
// the tile.  The minimum is taken with the reference's strict comparison,
// so a degenerate segment (NaN distance) is skipped as it is there.
__global__ __launch_bounds__(CSU_BLOCK) void k_wall_dist(int nc, const double* __restrict__ xc,
                                                         const double* __restrict__ yc, int nw,
                                                         const double* __restrict__ seg, double* __restrict__ dist) {
  __shared__ double sp[4][CSU_BLOCK];
  const int t = threadIdx.x, i = blockIdx.x * blockDim.x + t;
  const double x = i < nc ? xc[i] : 0.0, y = i < nc ? yc[i] : 0.0;
  double best = 1e6;
  for (int base = 0; base < nw; base += CSU_BLOCK) {
    const int w = base + t;
    if (w < nw) {
#pragma unroll
      for (int k = 0; k < 4; k++) sp[k][t] = seg[(size_t)k * nw + w];
    }
    __syncthreads();
    const int n = min(CSU_BLOCK, nw - base);
    for (int j = 0; j < n; j++) {
      const double p0x = sp[0][j], p0y = sp[1][j], p1x = sp[2][j], p1y = sp[3][j];
      double num = (p1x - p0x) * (x - p0x) + (p1y - p0y) * (y - p0y);
      double den = (p1x - p0x) * (p1x - p0x) + (p1y - p0y) * (p1y - p0y);
      double tt = num / den;
      if (tt > 1.) tt = 1.;
      else if (tt < 0.) tt = 0.;
      double dx = (p1x - p0x) * tt + p0x - x;
      double dy = (p1y - p0y) * tt + p0y - y;
      double dd = sqrt(dx * dx + dy * dy);
      if (best > dd) best = dd;
    }
    __syncthreads();
  }
  if (i < nc) dist[i] = best;
}

}  // namespace csu
