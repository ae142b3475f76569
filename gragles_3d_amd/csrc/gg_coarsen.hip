// Grid coarsening: grainhdl::gridCoarsement's coarsening branch
// (src/grainhdl.cpp:1446-1492), the per-grain LSbox::resizeGrid
// (src/box.cpp:2595-2697).  The IDLocal recalculation and the contour that
// follow it reuse the comparison and level-set kernels.
//
// Work per grain is one trilinear interpolation per voxel of the new box (8
// reads of the old box, 1 write); a coarsening runs once in many timesteps, so
// the kernel is a plain workgroup-per-grain loop.  Every arithmetic step is
// written in the reference's order (and the library is built without FMA
// contraction), so FP64 results are bit-identical to the CPU oracle's.
#include <cmath>

#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kCT = 256;

// m_new{X,Y,Z}{Min,Max} = (old - gb) * (h_old/h_new) + 0.5 + gb, truncated to
// int (src/box.cpp:2600-2611); the input buffer takes them cubed on the new
// grid (src/box.cpp:2618-2620).
__global__ __launch_bounds__(kCT) void k_coarsen_bounds(SimParams sp, CoarsenArgs a,
                                                        const int32_t* __restrict__ state,
                                                        const int32_t* __restrict__ src_b, int32_t* dst_b,
                                                        int32_t* nb) {
  for (int g = 1 + blockIdx.x * blockDim.x + threadIdx.x; g <= sp.N; g += gridDim.x * blockDim.x) {
    if (state[g] != kAlive) continue;
    int b[6];
    for (int k = 0; k < 6; ++k) {
      b[k] = (int)((src_b[6 * g + k] - a.gb) * a.up + 0.5 + a.gb);
      nb[6 * g + k] = b[k];
    }
    cube_bounds(b, sp.ngp);
    for (int k = 0; k < 6; ++k) dst_b[6 * g + k] = b[k];
  }
}

// One workgroup per grain: every voxel (j, i, k) of the new box samples the old
// box at ((j, i, k) - gb) * (h_new/h_old) + gb (src/box.cpp:2622-2690).  The
// bounds test keeps the interpolation cell inside the old box less two layers
// at its upper faces; outside it the value is -delta.
template <class T>
__global__ __launch_bounds__(kCT) void k_coarsen_resample(SimParams sp, CoarsenArgs a,
                                                          const int32_t* __restrict__ state, DescDev src,
                                                          DescDev dst) {
  for (int g = 1 + blockIdx.x; g <= sp.N; g += gridDim.x) {
    if (state[g] != kAlive) continue;
    const int32_t* sb = src.bounds + 6 * g;
    const int32_t* db = dst.bounds + 6 * g;
    const int sx0 = sb[0], sy0 = sb[1], sz0 = sb[2], sx1 = sb[3], sy1 = sb[4], sz1 = sb[5];
    const int snx = sx1 - sx0, sny = sy1 - sy0;
    const int dx0 = db[0], dy0 = db[1], dz0 = db[2];
    const int nx = db[3] - dx0, ny = db[4] - dy0, nz = db[5] - dz0;
    const T* S = static_cast<const T*>(src.arena) + src.off[g];
    T* D = static_cast<T*>(dst.arena) + dst.off[g];
    const int gb = a.gb;
    // getValueAt(y + gb, x + gb, z + gb) of the old box (x fastest)
    auto V = [&](double y, double x, double z) -> double {
      const int yy = (int)(y + gb) - sy0, xx = (int)(x + gb) - sx0, zz = (int)(z + gb) - sz0;
      return (double)S[((int64_t)zz * sny + yy) * snx + xx];
    };
    const int total = nx * ny * nz;  // at most 96^3
    for (int v = threadIdx.x; v < total; v += kCT) {
      const int q = v / nx;
      const int j = dx0 + (v - q * nx);
      const int i = dy0 + q % ny;
      const int k = dz0 + q / ny;
      const double px = (j - gb) * a.down, py = (i - gb) * a.down, pz = (k - gb) * a.down;
      const double xl = floor(px), xr = floor(px + 1);
      const double yu = floor(py), yo = floor(py + 1);
      const double zu = floor(pz), zo = floor(pz + 1);
      if (xr + gb > sx1 - 2 || yo + gb > sy1 - 2 || zo + gb > sz1 - 2 || yu + gb < sy0 || xl + gb < sx0 ||
          zu + gb < sz0) {
        D[v] = (T)(-a.delta);
        continue;
      }
      double ro = (xr - px) * V(yo, xl, zu) + (px - xl) * V(yo, xr, zu);
      double ru = (xr - px) * V(yu, xl, zu) + (px - xl) * V(yu, xr, zu);
      const double d_zu = (yo - py) * ru + (py - yu) * ro;
      ro = (xr - px) * V(yo, xl, zo) + (px - xl) * V(yo, xr, zo);
      ru = (xr - px) * V(yu, xl, zo) + (px - xl) * V(yu, xr, zo);
      const double d_zo = (yo - py) * ru + (py - yu) * ro;
      D[v] = (T)((zo - pz) * d_zu + (pz - zu) * d_zo);
    }
  }
}

}  // namespace

void launch_coarsen_bounds(const Launch& L, const SimParams& sp, const CoarsenArgs& a, const int32_t* state,
                           const int32_t* src_bounds, int32_t* dst_bounds, int32_t* nb) {
  const int blocks = (sp.N + kCT - 1) / kCT;
  hipLaunchKernelGGL(k_coarsen_bounds, dim3(blocks > 0 ? blocks : 1), dim3(kCT), 0, L.stream, sp, a, state,
                     src_bounds, dst_bounds, nb);
}

void launch_coarsen_resample(const Launch& L, const SimParams& sp, const CoarsenArgs& a, const int32_t* state,
                             const DescDev& src, const DescDev& dst, bool fp32) {
  const int grid = sp.N < 65536 ? (sp.N > 0 ? sp.N : 1) : 65536;
  if (fp32)
    hipLaunchKernelGGL(k_coarsen_resample<float>, dim3(grid), dim3(kCT), 0, L.stream, sp, a, state, src, dst);
  else
    hipLaunchKernelGGL(k_coarsen_resample<double>, dim3(grid), dim3(kCT), 0, L.stream, sp, a, state, src, dst);
}

}  // namespace gg
