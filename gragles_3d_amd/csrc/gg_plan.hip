// Work plan, arena layout and per-grain bookkeeping kernels.
//
// The plan groups the alive grains by box-size class and gives each class
// exclusive prefixes of three tile counts (z-layers, 64-column blocks and
// 32-column spectral blocks).  The same single-workgroup kernel lays out the
// boxes of one distance role back to back in its arena (the device form of the
// reference's per-LSbox DimensionalBuffer::resize, src/dimensionalBuffer.h:120-131)
// and checks the arena capacities.
#include "gg_internal.h"

namespace gg {

namespace {

constexpr int kPlanThreads = 1024;

// Exclusive workgroup scan of one per-thread quantity: wavefront scan by
// shuffles, then one exchange of the wavefront totals through LDS (two
// barriers, instead of a barrier ladder over the workgroup).
__device__ int64_t block_scan_excl(int64_t v, int64_t* sw, int64_t* total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int64_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) sw[wv] = x;
  __syncthreads();
  int64_t pre = 0, t = 0;
  for (int w = 0; w < kPlanThreads / 64; ++w) {
    const int64_t y = sw[w];
    if (w < wv) pre += y;
    t += y;
  }
  *total = t;
  __syncthreads();
  return pre + x - v;
}

__device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Side and size class of every grain, per position in the spatial order
// (one thread per grain), so the single-workgroup plan below reads one word
// per grain instead of a chain of dependent loads.
__global__ __launch_bounds__(256) void k_plan_classes(SimParams sp, const int32_t* __restrict__ state,
                                                      const int32_t* __restrict__ bounds, PlanDev P, int32_t* err) {
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < sp.N; idx += gridDim.x * blockDim.x) {
    const int g = P.order[idx];
    P.grain_pos[g] = -1;
    int v = -1;
    if (state[g] == kAlive) {
      const int32_t* b = bounds + 6 * g;
      const int n = b[3] - b[0];
      const int c = size_class(n);
      if (n < 1 || b[4] - b[1] != n || b[5] - b[2] != n || c < 0) atomicOr(err, kErrBoxSize);
      else v = n << 8 | c;
    }
    P.ncls[idx] = v;
  }
}

// The volume-sorted work queue: the grains of the spatial order sorted by box
// side, largest first (not-alive grains last), equal sides kept in the spatial
// order.  One workgroup: a side histogram, its descending exclusive scan, then
// rounds of kPlanThreads positions in which each lane's rank among equal sides
// comes from ballots inside its wavefront and the counts of the wavefronts
// before it, so the sort is stable and deterministic.
__global__ __launch_bounds__(kPlanThreads) void k_plan_vorder(SimParams sp, const int32_t* __restrict__ state,
                                                              const int32_t* __restrict__ bounds, PlanDev P) {
  constexpr int kBins = kMaxN + 1;  // side 0: not planned
  constexpr int kWaves = kPlanThreads / 64;
  __shared__ int s_start[kBins];
  __shared__ int s_wcnt[kWaves][kBins];
  const int N = sp.N, t = threadIdx.x, lane = t & 63, wv = t >> 6;
  auto key_at = [&](int idx) {
    const int g = P.order[idx];
    if (state[g] != kAlive) return 0;
    const int n = bounds[6 * g + 3] - bounds[6 * g];
    return n < 1 ? 0 : (n > kMaxN ? kMaxN : n);
  };
  for (int i = t; i < kBins; i += kPlanThreads) s_start[i] = 0;
  __syncthreads();
  for (int idx = t; idx < N; idx += kPlanThreads) atomicAdd(&s_start[key_at(idx)], 1);
  __syncthreads();
  if (t == 0) {
    int acc = 0;
    for (int k = kMaxN; k >= 0; --k) {
      const int c = s_start[k];
      s_start[k] = acc;
      acc += c;
    }
  }
  __syncthreads();
  for (int base = 0; base < N; base += kPlanThreads) {
    const int idx = base + t;
    const int k = idx < N ? key_at(idx) : -1;
    for (int i = t; i < kWaves * kBins; i += kPlanThreads) (&s_wcnt[0][0])[i] = 0;
    __syncthreads();
    int rank = 0;
    bool done = k < 0;
    for (;;) {  // one round per distinct side in the wavefront
      const unsigned long long pend = __ballot(!done);
      if (pend == 0ull) break;
      const int k0 = __shfl(k, __ffsll((long long)pend) - 1);
      const bool mine = !done && k == k0;
      const unsigned long long m = __ballot(mine);
      if (mine) {
        rank = __popcll(m & ((1ull << lane) - 1ull));
        done = true;
      }
      if (lane == 0) s_wcnt[wv][k0] = __popcll(m);
    }
    __syncthreads();
    if (k >= 0) {
      int pos = s_start[k] + rank;
      for (int w = 0; w < wv; ++w) pos += s_wcnt[w][k];
      P.vorder[pos] = P.order[idx];
    }
    __syncthreads();
    for (int i = t; i < kBins; i += kPlanThreads) {
      int add = 0;
      for (int w = 0; w < kWaves; ++w) add += s_wcnt[w][i];
      s_start[i] += add;
    }
    __syncthreads();
  }
}

// Exclusive scan of K values per thread over the workgroup, in place, all K
// in one pass (one barrier).  sw[w][k] keeps wave w's sum of value k, so the
// totals can be read from it afterwards (plan_total).
template <int K>
__device__ inline void block_scan_excl_multi(int (&x)[K], int (*sw)[K]) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    int s = x[k];
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(s, off);
      if (lane >= off) s += y;
    }
    if (lane == 63) sw[wv][k] = s;
    x[k] = s - x[k];
  }
  __syncthreads();
  for (int w = 0; w < wv; ++w) {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] += sw[w][k];
  }
}

template <int K>
__device__ inline int plan_total(int (*sw)[K], int k) {
  int t = 0;
  for (int w = 0; w < kPlanThreads / 64; ++w) t += sw[w][k];
  return t;
}

__global__ __launch_bounds__(kPlanThreads) void k_plan(SimParams sp, const int32_t* __restrict__ state,
                                                       const int32_t* __restrict__ bounds, int64_t* off_out,
                                                       int64_t* spec_off, PlanDev P, int64_t arena_cap,
                                                       int64_t spec_cap, int32_t* err, int64_t* voxels) {
  const int N = sp.N;
  const int t = threadIdx.x;
  const int chunk = (N + kPlanThreads - 1) / kPlanThreads;
  const int gb = 1 + t * chunk;
  const int ge = min(N + 1, gb + chunk);

  // per thread: [0..4) grains per class, [4 + q * 4 + c] tiles of quantity q in
  // class c (32-bit, as the plan stores them), then voxels and spectrum (64-bit).
  // A grain's class selects its slots by unrolled compares, never as an index,
  // so the counters stay in registers.
  constexpr int K = 4 + kNumQ * kClasses;
  __shared__ int64_t sw[kPlanThreads / 64];
  __shared__ int swk[kPlanThreads / 64][K];
  __shared__ int s_tot[K];
  int loc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) loc[k] = 0;
  int64_t lvox = 0, lspec = 0;
  // grains are visited in the spatial order, so class lists, tiles and the
  // arena layout keep neighbouring boxes close together
  for (int idx = gb; idx < ge; ++idx) {
    const int v = P.ncls[idx - 1];
    if (v < 0) continue;
    const int n = v >> 8, c = v & 0xFF;
    const int tc = (int)ceil_div((int64_t)n * n, kColBlock), ts = (int)ceil_div((int64_t)n * (n / 2 + 1), kSpecBlock);
#pragma unroll
    for (int cc = 0; cc < kClasses; ++cc) {
      const bool h = c == cc;
      loc[cc] += h ? 1 : 0;
      loc[4 + kQLayers * kClasses + cc] += h ? n : 0;
      loc[4 + kQCols * kClasses + cc] += h ? tc : 0;
      loc[4 + kQSpecCols * kClasses + cc] += h ? ts : 0;
    }
    lvox += (int64_t)n * n * n;
    lspec += (int64_t)n * n * (n / 2 + 1);
  }
  int* const ex = loc;  // becomes the exclusive prefixes
  block_scan_excl_multi<K>(loc, swk);
  if (t < K) s_tot[t] = plan_total<K>(swk, t);  // published by the barriers of the scans below
  int64_t tvox, tspec;
  int64_t exv = block_scan_excl(lvox, sw, &tvox);
  int64_t exs = block_scan_excl(lspec, sw, &tspec);

  int cls_start[kClasses];
  {
    int acc = 0;
#pragma unroll
    for (int c = 0; c < kClasses; ++c) {
      cls_start[c] = acc;
      acc += s_tot[c];
    }
  }
  // A layout that does not fit its arena leaves an empty plan, so no kernel of
  // this step touches memory beyond the arenas; the host reports the error bit.
  const bool overflow = (off_out && tvox > arena_cap) || (spec_off && tspec > spec_cap);
  if (t == 0) {
    const int* tot = s_tot;
    int64_t lb = 0;
#pragma unroll
    for (int c = 0; c < kClasses; ++c) {
      P.cls_count[c] = overflow ? 0 : tot[c];
      P.cls_start[c] = cls_start[c];
#pragma unroll
      for (int q = 0; q < kNumQ; ++q) P.totals[q * kClasses + c] = overflow ? 0 : tot[4 + q * kClasses + c];
      P.layer_base[c] = lb;
      lb += tot[4 + kQLayers * kClasses + c];
#pragma unroll
      for (int q = 0; q < kNumQ; ++q) P.pre[q][cls_start[c] + c + tot[c]] = tot[4 + q * kClasses + c];
    }
#pragma unroll
    for (int q = 0; q < kNumQ; ++q) {
      int tb = 0;
#pragma unroll
      for (int c = 0; c < kClasses; ++c) {
        P.tile_base[q * kClasses + c] = tb;
        tb += overflow ? 0 : tot[4 + q * kClasses + c];
      }
    }
    if (voxels) *voxels = tvox;
    if (off_out && tvox > arena_cap) atomicOr(err, kErrArena);
    if (spec_off && tspec > spec_cap) atomicOr(err, kErrSpec);
  }
  for (int idx = gb; idx < ge; ++idx) {
    const int v = P.ncls[idx - 1];
    if (v < 0) continue;
    const int n = v >> 8, c = v & 0xFF;
    const int g = P.order[idx - 1];
    const int tc = (int)ceil_div((int64_t)n * n, kColBlock), ts = (int)ceil_div((int64_t)n * (n / 2 + 1), kSpecBlock);
    int pos = 0, el = 0, ec = 0, es = 0;
#pragma unroll
    for (int cc = 0; cc < kClasses; ++cc) {
      const bool h = c == cc;
      pos = h ? cls_start[cc] + ex[cc] : pos;
      el = h ? ex[4 + kQLayers * kClasses + cc] : el;
      ec = h ? ex[4 + kQCols * kClasses + cc] : ec;
      es = h ? ex[4 + kQSpecCols * kClasses + cc] : es;
    }
    P.grains[pos] = g;
    P.grain_pos[g] = pos;
    const int seg = pos + c;  // class c's segment of the tile prefixes has one extra slot per lower class
    P.pre[kQLayers][seg] = el;
    P.pre[kQCols][seg] = ec;
    P.pre[kQSpecCols][seg] = es;
#pragma unroll
    for (int cc = 0; cc < kClasses; ++cc) {
      const bool h = c == cc;
      ex[cc] += h ? 1 : 0;
      ex[4 + kQLayers * kClasses + cc] += h ? n : 0;
      ex[4 + kQCols * kClasses + cc] += h ? tc : 0;
      ex[4 + kQSpecCols * kClasses + cc] += h ? ts : 0;
    }
    if (off_out) off_out[g] = exv;
    if (spec_off) spec_off[g] = exs;
    exv += (int64_t)n * n * n;
    exs += (int64_t)n * n * (n / 2 + 1);
  }
}

// Tile map of the plan: for every alive grain and quantity, one entry per tile
// (grain id << 8 | tile index inside the grain), so a kernel finds its tile's
// grain with one load.
__global__ __launch_bounds__(256) void k_tilemap(SimParams sp, PlanDev P) {
  for (int pos = blockIdx.x * blockDim.x + threadIdx.x; pos < sp.N; pos += gridDim.x * blockDim.x) {
    int c = 0;
    while (c < kClasses && !(pos >= P.cls_start[c] && pos < P.cls_start[c] + P.cls_count[c])) ++c;
    if (c == kClasses) continue;
    const uint32_t g = (uint32_t)P.grains[pos];
    const int seg = pos + c;  // P.pre segment of class c starts at cls_start[c] + c
    for (int q = 0; q < kNumQ; ++q) {
      const int b0 = P.pre[q][seg], b1 = P.pre[q][seg + 1];
      uint32_t* m = P.tmap[q] + P.tile_base[q * kClasses + c];
      for (int l = b0; l < b1; ++l) m[l] = g << 8 | (uint32_t)(l - b0);
    }
  }
}

// calculateDistanceFunction (src/box.cpp:374-397): +h inside the grain's
// labelled voxels, -h elsewhere and on the border shell.  Only the grid layers
// k in [k_lo, k_hi) are written; idfield holds those layers (idfield[0] is
// voxel (0, 0, k_lo)).
template <class T>
__global__ __launch_bounds__(256) void k_init_phi(SimParams sp, PlanDev P, DescDev din,
                                                  const int32_t* __restrict__ idfield, int k_lo, int k_hi) {
  const int total = plan_total(P, kQLayers);
  T* arena = (T*)din.arena;
  const int L = sp.ngp, lo = sp.gb, hi = sp.ngp - sp.gb - 1;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int c, g, kl;
    plan_tile_any(P, kQLayers, tile, c, g, kl);
    const int32_t* b = din.bounds + 6 * g;
    const int n = b[3] - b[0];
    const int k = b[2] + kl;
    if (k < k_lo || k >= k_hi) continue;
    T* dst = arena + din.off[g] + (int64_t)kl * n * n;
    for (int v = threadIdx.x; v < n * n; v += blockDim.x) {
      int i = b[1] + v / n, j = b[0] + v % n;
      T val;
      if (i < lo || i > hi || j < lo || j > hi || k < lo || k > hi)
        val = (T)(-sp.h);
      else if (idfield[((int64_t)(k - k_lo) * L + i) * L + j] == g)
        val = (T)sp.h;
      else
        val = (T)(-sp.h);
      dst[v] = val;
    }
  }
}

// resizeIDLocalToDistanceBuffer (src/box.cpp:559-567): the IDLocal buffer the
// comparison writes takes the output buffer's bounds and layout.
// Vanished grains follow too: one that vanished at the last level set after
// being planned is still listed by the current plan (the comparison then
// writes its IDLocal, which must land in its own planned region).
__global__ void k_idl_resize(SimParams sp, const int32_t* __restrict__ state, DescDev dout, IdlDev il) {
  for (int g = 1 + blockIdx.x * blockDim.x + threadIdx.x; g <= sp.N; g += gridDim.x * blockDim.x) {
    if (state[g] != kAlive && state[g] != kVanished) continue;
    il.off[g] = dout.off[g];
    for (int a = 0; a < 6; ++a) il.bounds[6 * g + a] = dout.bounds[6 * g + a];
  }
}

// End of executeRedistancing (src/box.cpp:2582-2585): the input buffer takes
// the output buffer's bounds (and, here, its arena layout).
// A grain that vanished at this level set takes them too: it was planned
// (and redistanced) when its hull was still to come, and the plan that becomes
// the current one lists it, so its input boxes must be its planned ones.
__global__ void k_post_redist(SimParams sp, const int32_t* __restrict__ state, DescDev din, DescDev dout) {
  for (int g = 1 + blockIdx.x * blockDim.x + threadIdx.x; g <= sp.N; g += gridDim.x * blockDim.x) {
    if (state[g] != kAlive && state[g] != kVanished) continue;
    din.off[g] = dout.off[g];
    for (int a = 0; a < 6; ++a) din.bounds[6 * g + a] = dout.bounds[6 * g + a];
  }
}

// countGrains (src/grainhdl.cpp:936-943): grains not deleted; appended to the
// on-device trajectory.
__global__ __launch_bounds__(1024) void k_count(SimParams sp, const int32_t* __restrict__ state, int32_t* step,
                                                int32_t* hist, int32_t cap, int32_t* err) {
  __shared__ int s[1024];
  int c = 0;
  for (int g = 1 + threadIdx.x; g <= sp.N; g += 1024) c += state[g] != kDeleted;
  s[threadIdx.x] = c;
  __syncthreads();
  for (int d = 512; d > 0; d >>= 1) {
    if (threadIdx.x < d) s[threadIdx.x] += s[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int st = *step;
    if (st < cap) hist[st] = s[0];
    else atomicOr(err, kErrHistory);
    *step = st + 1;
  }
}

}  // namespace

void launch_plan(const Launch& L, const SimParams& sp, const int32_t* state, const int32_t* bounds,
                 int64_t* off_out, int64_t* spec_off, const PlanDev& P, int64_t arena_cap, int64_t spec_cap,
                 int32_t* err, int64_t* voxels) {
  // with a volume-sorted queue the plan visits the grains in that order
  PlanDev Q = P;
  if (P.vorder) {
    hipLaunchKernelGGL(k_plan_vorder, dim3(1), dim3(kPlanThreads), 0, L.stream, sp, state, bounds, P);
    Q.order = P.vorder;
  }
  const int cblocks = (sp.N + 255) / 256 < 4096 ? (sp.N + 255) / 256 : 4096;
  hipLaunchKernelGGL(k_plan_classes, dim3(cblocks > 0 ? cblocks : 1), dim3(256), 0, L.stream, sp, state, bounds, Q,
                     err);
  hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, L.stream, sp, state, bounds, off_out, spec_off, Q,
                     arena_cap, spec_cap, err, voxels);
  const int blocks = (sp.N + 255) / 256 < 1024 ? (sp.N + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_tilemap, dim3(blocks), dim3(256), 0, L.stream, sp, Q);
}

void launch_init_phi(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din,
                     const int32_t* idfield, bool fp32, int k_lo, int k_hi) {
  if (fp32)
    hipLaunchKernelGGL(k_init_phi<float>, dim3(L.grid), dim3(256), 0, L.stream, sp, P, din, idfield, k_lo, k_hi);
  else
    hipLaunchKernelGGL(k_init_phi<double>, dim3(L.grid), dim3(256), 0, L.stream, sp, P, din, idfield, k_lo, k_hi);
}

void launch_idl_resize(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& dout,
                       const IdlDev& il) {
  int blocks = (sp.N + 255) / 256;
  hipLaunchKernelGGL(k_idl_resize, dim3(blocks), dim3(256), 0, L.stream, sp, state, dout, il);
}

void launch_post_redist(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                        const DescDev& dout) {
  int blocks = (sp.N + 255) / 256;
  hipLaunchKernelGGL(k_post_redist, dim3(blocks), dim3(256), 0, L.stream, sp, state, din, dout);
}

void launch_count(const Launch& L, const SimParams& sp, const int32_t* state, int32_t* step, int32_t* history,
                  int32_t history_cap, int32_t* err) {
  hipLaunchKernelGGL(k_count, dim3(1), dim3(1024), 0, L.stream, sp, state, step, history, history_cap, err);
}

}  // namespace gg
