// Comparison step of every grain: executeComparison (src/box.cpp:891-966) +
// BoundaryIntersection (:968-990) + boundaryCondition (:997-1088) +
// executeSetComparison (:814-878), fused per (grain, z-layer) tile, one launch
// per box-size class.
//
// Per tile the workgroup first compacts, in comparison-list order, the
// neighbours whose input box covers this layer and overlaps the grain's box
// in x and y.  Each lane owns a fixed set of voxels of the layer (registers)
// and walks the compacted neighbours; for one neighbour all its voxel loads
// are independent, so they are in flight together.  Per voxel the updates
// still happen in list order with a strict '>' (ties keep the earlier
// neighbour), which is the voxel-wise result of the reference's
// neighbour-outer loop.  Then the wall distance, IDLocal (new id, or the
// previous step's id at the same linear index when nothing covers the voxel),
// 0.5*(in - max) inside 0.7*delta, and the bounding box of phi >= 0.
#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kMaxAct = 1024;  // >= the largest comparison-list capacity gg_create accepts
constexpr int kCT = 256;

struct ActNb {
  int32_t id;
  int16_t x0, x1, y0, y1;
  int32_t nn;
  int64_t base;
};

// Value of boundaryCondition's `dist` at voxel (j,i,k): the reference assigns
// it in every branch except when all three axis distances are >= 0 and one is
// 0; there it keeps the value of the previous voxel in its (k,i,j) loop over
// the box (starting from 0).  Walk back in that order to the last assignment.
__device__ double bc_dist(int j, int i, int k, int x0, int y0, int z0, int x1, int y1, int gb, int m) {
  for (;;) {
    const double dXMin = -(j - gb + 1), dYMin = -(i - gb + 1), dZMin = -(k - gb + 1);
    const double dXMax = (j - (m - 1 - gb)), dYMax = (i - (m - 1 - gb)), dZMax = (k - (m - 1 - gb));
    const double dX = fabs(dXMin) < fabs(dXMax) ? dXMin : dXMax;
    const double dY = fabs(dYMin) < fabs(dYMax) ? dYMin : dYMax;
    const double dZ = fabs(dZMin) < fabs(dZMax) ? dZMin : dZMax;
    if (dX > 0 && dY > 0 && dZ > 0) return sqrt((double)dX * dX + dY * dY + dZ * dZ);
    if (dX < 0 && dY < 0 && dZ < 0) return fmax(fmax(dX, dY), dZ);
    if (dX >= 0) {
      if (dY < 0 && dZ < 0) return dX;
      if (dY < 0 && dZ >= 0) return sqrt((double)dX * dX + dZ * dZ);
      if (dY >= 0 && dZ < 0) return sqrt((double)dX * dX + dY * dY);
    } else if (dY >= 0) {
      if (dZ < 0 && dX < 0) return dY;
      if (dX < 0 && dZ >= 0) return sqrt((double)dY * dY + dZ * dZ);
    } else if (dZ >= 0) {
      if (dY < 0 && dX < 0) return dZ;
    }
    // carry: previous voxel in loop order
    if (j > x0) {
      --j;
    } else {
      j = x1 - 1;
      if (i > y0) {
        --i;
      } else {
        i = y1 - 1;
        if (k > z0) --k;
        else return 0.0;
      }
    }
  }
}

template <class T, int NMAX>
__global__ __launch_bounds__(kCT) void k_comparison(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                    ListDev ld, IdlDev ilo, IdlDev iln, int32_t* newb) {
  constexpr int VPT = (NMAX * NMAX + kCT - 1) / kCT;  // voxels of one layer per lane
  __shared__ ActNb act[kMaxAct];
  __shared__ int s_wave[kCT / 64];
  __shared__ int s_red[6][kCT / 64];
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const T lim = (T)(0.7 * sp.delta);
  const int maxL = ld.max_list;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4];
    const int n = x1 - x0, nn2 = n * n;
    const int k = z0 + kl;
    const int lsel = ld.sel[g] ^ 1;  // m_comparisonList
    const int32_t* lst = ld.buf[lsel] + (int64_t)g * maxL;
    const int len = ld.len[lsel][g];
    // compact the neighbours covering this layer, keeping list order
    int cnt = 0;
    for (int base = 0; base < len; base += kCT) {
      const int idx = base + tid;
      bool keep = false;
      ActNb e;
      if (idx < len) {
        const int nb = lst[idx];
        const int32_t* q = din.bounds + 6 * nb;
        if (q[2] <= k && k < q[5] && max(x0, q[0]) < min(x1, q[3]) && max(y0, q[1]) < min(y1, q[4])) {
          keep = true;
          e.id = nb;
          e.x0 = (int16_t)q[0]; e.x1 = (int16_t)q[3]; e.y0 = (int16_t)q[1]; e.y1 = (int16_t)q[4];
          e.nn = q[3] - q[0];
          e.base = din.off[nb] + (int64_t)(k - q[2]) * e.nn * e.nn;
        }
      }
      const uint64_t m = __ballot(keep);
      if (lane == 0) s_wave[wv] = __popcll(m);
      __syncthreads();
      int before = cnt;
      for (int w = 0; w < wv; ++w) before += s_wave[w];
      const int pos = before + __popcll(m & ((1ull << lane) - 1ull));
      if (keep && pos < kMaxAct) act[pos] = e;
      for (int w = 0; w < kCT / 64; ++w) cnt += s_wave[w];
      __syncthreads();
    }
    if (cnt > kMaxAct) cnt = kMaxAct;

    // The lane's voxels: wavefront wv owns the VPT consecutive 64-voxel segments
    // starting at wv*VPT*64, i.e. a band of rows, so most neighbours miss most
    // segments and are skipped with one uniform test.  Packed (i - y0) << 16 | (j - x0).
    int ij[VPT];
    T best[VPT];
    uint32_t bid[VPT];
    int seg_lo[VPT], seg_hi[VPT];  // local row range of each segment (uniform)
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      const int s0 = (wv * VPT + r) * 64;
      const int v = s0 + lane;
      ij[r] = v < nn2 ? ((v / n) << 16) | (v % n) : -1;
      best[r] = (T)(-1.0);
      bid[r] = 0xFFFFFFFFu;  // not covered
      seg_lo[r] = s0 < nn2 ? s0 / n : n;
      seg_hi[r] = min(s0 + 63, nn2 - 1) / n;
    }
    for (int a = 0; a < cnt; ++a) {
      const ActNb e = act[a];
      const int ex0 = e.x0 - x0, ex1 = e.x1 - x0, ey0 = e.y0 - y0, ey1 = e.y1 - y0;
      const int64_t eb = e.base - (int64_t)ey0 * e.nn - ex0;  // value at local (il, jl): eb + il*nn + jl
      // all loads of this neighbour first, then the ordered updates
      T val[VPT];
      bool cov[VPT];
#pragma unroll
      for (int r = 0; r < VPT; ++r) {
        cov[r] = false;
        if (seg_lo[r] < ey1 && seg_hi[r] >= ey0) {
          const int il = ij[r] >> 16, jl = ij[r] & 0xFFFF;
          cov[r] = ij[r] >= 0 && jl >= ex0 && jl < ex1 && il >= ey0 && il < ey1;
          if (cov[r]) val[r] = ain[eb + (int64_t)il * e.nn + jl];
        }
      }
#pragma unroll
      for (int r = 0; r < VPT; ++r) {
        if (cov[r] && val[r] > best[r]) {
          best[r] = val[r];
          bid[r] = (uint32_t)e.id;
        }
      }
    }
    // BoundaryIntersection on the output box (equal to the input box here)
    const int32_t* ob = dout.bounds + 6 * g;
    const int lo = sp.gb + sp.tube, hi = sp.ngp - sp.gb - sp.tube;
    const bool wall = !(ob[0] > lo && ob[3] < hi && ob[1] > lo && ob[4] < hi && ob[2] > lo && ob[5] < hi);
    const int32_t* olb = ilo.bounds + 6 * g;
    const int64_t oldsize = (int64_t)(olb[3] - olb[0]) * (olb[4] - olb[1]) * (olb[5] - olb[2]);
    const uint32_t* __restrict__ idl_old = ilo.arena + ilo.off[g];
    uint32_t* __restrict__ idl_new = iln.arena + iln.off[g];
    const int64_t lbase = (int64_t)kl * nn2;
    const T* own = ain + din.off[g] + lbase;
    T* dst = aout + dout.off[g] + lbase;
    // maxima start at -INT_MAX so that their negation (reduced with min below) cannot overflow
    int bx0 = INT_MAX, by0 = INT_MAX, bz0 = INT_MAX, bx1 = -INT_MAX, by1 = -INT_MAX, bz1 = -INT_MAX;
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      if (ij[r] < 0) continue;
      const int il = ij[r] >> 16, jl = ij[r] & 0xFFFF;
      const int i = y0 + il, j = x0 + jl;
      const int v = il * n + jl;
      if (wall) {
        const double dist = bc_dist(j, i, k, x0, y0, z0, x1, y1, sp.gb, sp.ngp);
        if (dist > -sp.gb) {
          const T dh = (T)(dist * sp.h);
          if (dh > best[r]) {
            best[r] = dh;
            bid[r] = 0;
          }
        }
      }
      const int64_t li = lbase + v;
      uint32_t idv = bid[r];
      if (idv == 0xFFFFFFFFu) idv = li < oldsize ? idl_old[li] : 0u;
      idl_new[li] = idv;
      const T a_in = own[v];
      const T o = (a_in < T(0) ? -a_in : a_in) < lim ? (T)0.5 * (a_in - best[r]) : a_in;
      dst[v] = o;
      if (o >= T(0)) {
        bx0 = min(bx0, j); bx1 = max(bx1, j);
        by0 = min(by0, i); by1 = max(by1, i);
        bz0 = min(bz0, k); bz1 = max(bz1, k);
      }
    }
    // wave then workgroup min/max of the non-negative voxels
    int vals[6] = {bx0, by0, bz0, -bx1, -by1, -bz1};
    for (int q = 0; q < 6; ++q) {
      int x = vals[q];
      for (int off = 32; off > 0; off >>= 1) x = min(x, __shfl_xor(x, off));
      if (lane == 0) s_red[q][wv] = x;
    }
    __syncthreads();
    if (tid == 0) {
      int rr[6];
      for (int q = 0; q < 6; ++q) {
        rr[q] = s_red[q][0];
        for (int w = 1; w < kCT / 64; ++w) rr[q] = min(rr[q], s_red[q][w]);
      }
      if (rr[0] != INT_MAX) {
        int32_t* nb = newb + 6 * g;
        atomicMin(&nb[0], rr[0]);
        atomicMin(&nb[1], rr[1]);
        atomicMin(&nb[2], rr[2]);
        atomicMax(&nb[3], -rr[3]);
        atomicMax(&nb[4], -rr[4]);
        atomicMax(&nb[5], -rr[5]);
      }
    }
    __syncthreads();
  }
}

template <class T>
void comparison_all(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din,
                    const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  hipLaunchKernelGGL((k_comparison<T, 32>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 0, din, dout, ld, ilo,
                     iln, newb);
  hipLaunchKernelGGL((k_comparison<T, 48>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 1, din, dout, ld, ilo,
                     iln, newb);
  hipLaunchKernelGGL((k_comparison<T, 64>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 2, din, dout, ld, ilo,
                     iln, newb);
  hipLaunchKernelGGL((k_comparison<T, 96>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 3, din, dout, ld, ilo,
                     iln, newb);
}

}  // namespace

void launch_comparison(const Launch& L, const SimParams& sp, const PlanDev& P, const int32_t* state,
                       const DescDev& din, const DescDev& dout, const ListDev& ld, const IdlDev& il_old,
                       const IdlDev& il_new, int32_t* newb, bool fp32) {
  (void)state;
  if (fp32) comparison_all<float>(L, sp, P, din, dout, ld, il_old, il_new, newb);
  else comparison_all<double>(L, sp, P, din, dout, ld, il_old, il_new, newb);
}

}  // namespace gg
