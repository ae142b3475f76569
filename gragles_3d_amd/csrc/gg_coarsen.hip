// Whole-grid work on the grain boxes: grid coarsening (grainhdl::gridCoarsement's
// coarsening branch, src/grainhdl.cpp:1446-1492, with the per-grain
// LSbox::resizeGrid, src/box.cpp:2595-2697; the IDLocal recalculation and the
// contour that follow it reuse the comparison and level-set kernels) and the
// voxel container of grainhdl::saveNetworkAsVoxelContainer
// (src/grainhdl.cpp:1074-1154).
//
// Work per grain is one trilinear interpolation per voxel of the new box (8
// reads of the old box, 1 write); a coarsening runs once in many timesteps, so
// the kernel is a plain workgroup-per-grain loop.  Every arithmetic step is
// written in the reference's order (and the library is built without FMA
// contraction), so FP64 results are bit-identical to the CPU oracle's.
#include <algorithm>
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

// saveNetworkAsVoxelContainer's copyDataToContainer (src/box.cpp:3008-3032):
// every voxel of an alive grain's output buffer with phi >= 0 inside
// [gb, gb + size)^3 names the grain in the size^3 container (x fastest).
// Where several grains do, the largest id stays: the reference's serial
// ascending-id order.  One workgroup per grain.
template <class T>
__global__ __launch_bounds__(kCT) void k_container(SimParams sp, const int32_t* __restrict__ state, DescDev d,
                                                   uint32_t* cont, int size) {
  for (int g = 1 + blockIdx.x; g <= sp.N; g += gridDim.x) {
    if (state[g] != kAlive) continue;
    const int32_t* b = d.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2];
    const int nx = b[3] - x0, ny = b[4] - y0, nz = b[5] - z0;
    const T* P = static_cast<const T*>(d.arena) + d.off[g];
    const int gb = sp.gb, total = nx * ny * nz;
    for (int v = threadIdx.x; v < total; v += kCT) {
      if (!(T(0) <= P[v])) continue;
      const int q = v / nx;
      const int j = x0 + (v - q * nx) - gb, i = y0 + q % ny - gb, k = z0 + q / ny - gb;
      if (i < 0 || j < 0 || k < 0 || i >= size || j >= size || k >= size) continue;
      atomicMax(&cont[((int64_t)k * size + i) * size + j], (uint32_t)g);
    }
  }
}

__global__ __launch_bounds__(kCT) void k_fill_u32(uint32_t* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)kCT + threadIdx.x; i < n; i += (int64_t)gridDim.x * kCT) p[i] = 0u;
}

}  // namespace

void launch_container(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& d, uint32_t* cont,
                      int size, bool fp32) {
  const int64_t n = (int64_t)size * size * size;
  const int fb = (int)std::min<int64_t>((n + kCT - 1) / kCT, 65536);
  hipLaunchKernelGGL(k_fill_u32, dim3(fb > 0 ? fb : 1), dim3(kCT), 0, L.stream, cont, n);
  const int grid = sp.N < 65536 ? (sp.N > 0 ? sp.N : 1) : 65536;
  if (fp32)
    hipLaunchKernelGGL(k_container<float>, dim3(grid), dim3(kCT), 0, L.stream, sp, state, d, cont, size);
  else
    hipLaunchKernelGGL(k_container<double>, dim3(grid), dim3(kCT), 0, L.stream, sp, state, d, cont, size);
}

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
