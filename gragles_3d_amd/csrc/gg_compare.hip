// Comparison lists and the comparison step of every grain.
//
// k_lists: LSbox::computeSecondOrderNeighbours (src/box.cpp:1090-1134) with
// checkIntersection (src/box.cpp:880-889).  One wavefront per grain; a bitmap
// of grain ids in LDS replaces the reference's linear duplicate scan, and the
// candidates are appended strictly in the reference's order (previous list
// entries that still exist, then neighbours-of-neighbours in traversal order,
// first occurrence wins), because that order decides comparison ties.
//
// k_comparison: executeComparison (src/box.cpp:891-966) + BoundaryIntersection
// (:968-990) + boundaryCondition (:997-1088) + executeSetComparison
// (:814-878), fused per (grain, z-layer) tile.  Per voxel the neighbour maximum
// is taken in comparison-list order with a strict '>' (ties keep the earlier
// neighbour), exactly the voxel-wise result of the reference's
// neighbour-outer loop.  The neighbours whose input box covers the layer are
// compacted into LDS first, so each voxel only tests those.
#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kMaxAct = 1024;  // >= the largest comparison-list capacity gg_create accepts

__global__ __launch_bounds__(64) void k_lists(SimParams sp, const int32_t* __restrict__ state, DescDev din,
                                              DescDev dout, ListDev ld, int32_t* newb, int32_t* err) {
  extern __shared__ uint32_t bitmap[];
  const int N = sp.N;
  const int lane = threadIdx.x;
  const int words = (N + 1 + 31) / 32;
  for (int w = lane; w < words; w += 64) bitmap[w] = 0u;
  __syncthreads();
  const int maxL = ld.max_list;
  for (int g = 1 + blockIdx.x; g <= N; g += gridDim.x) {
    if (state[g] != kAlive) continue;
    if (lane == 0) {
      // executeSetComparison's starting values m_newXMin = out.maxX ... (src/box.cpp:815-820)
      const int32_t* ob = dout.bounds + 6 * g;
      int32_t* nb = newb + 6 * g;
      nb[0] = ob[3]; nb[1] = ob[4]; nb[2] = ob[5];
      nb[3] = ob[0]; nb[4] = ob[1]; nb[5] = ob[2];
    }
    const int cur = ld.sel[g];
    const int32_t* src = ld.buf[cur] + (int64_t)g * maxL;
    const int srclen = ld.len[cur][g];
    int32_t* dst = ld.buf[cur ^ 1] + (int64_t)g * maxL;
    const int32_t* gbnd = din.bounds + 6 * g;
    const int gx0 = gbnd[0], gy0 = gbnd[1], gz0 = gbnd[2], gx1 = gbnd[3], gy1 = gbnd[4], gz1 = gbnd[5];
    int outlen = 0;
    // the grain itself and the wall never become candidates (src/box.cpp:1118-1122)
    if (lane == 0) {
      bitmap[g >> 5] |= 1u << (g & 31);
      bitmap[0] |= 1u;
    }
    __syncthreads();
    // previous neighbours that still exist, in order
    for (int base = 0; base < srclen; base += 64) {
      const int idx = base + lane;
      int id = 0;
      bool keep = false;
      if (idx < srclen) {
        id = src[idx];
        keep = state[id] == kAlive;
      }
      const uint64_t m = __ballot(keep);
      const int pos = outlen + __popcll(m & ((1ull << lane) - 1ull));
      if (keep) {
        if (pos < maxL) dst[pos] = id;
        atomicOr(&bitmap[id >> 5], 1u << (id & 31));
      }
      outlen += __popcll(m);
    }
    __syncthreads();
    // neighbours of neighbours whose input boxes intersect (inclusive bounds)
    for (int q = 0; q < srclen; ++q) {
      const int nbid = src[q];
      const int nsel = ld.sel[nbid];
      const int32_t* nsrc = ld.buf[nsel] + (int64_t)nbid * maxL;
      const int nlen = ld.len[nsel][nbid];
      for (int base = 0; base < nlen; base += 64) {
        const int idx = base + lane;
        int cid = 0;
        bool p = false;
        if (idx < nlen) {
          cid = nsrc[idx];
          if (state[cid] == kAlive) {
            const int32_t* cb = din.bounds + 6 * cid;
            p = !(gx0 > cb[3] || gx1 < cb[0] || gy0 > cb[4] || gy1 < cb[1] || gz0 > cb[5] || gz1 < cb[2]);
          }
        }
        uint64_t m = __ballot(p);
        while (m) {
          const int l = __ffsll((unsigned long long)m) - 1;
          m &= m - 1;
          const int c = __shfl(cid, l);
          const bool fresh = ((bitmap[c >> 5] >> (c & 31)) & 1u) == 0u;
          if (fresh) {
            if (lane == 0) {
              bitmap[c >> 5] |= 1u << (c & 31);
              if (outlen < maxL) dst[outlen] = c;
            }
            ++outlen;
          }
          __syncthreads();
        }
      }
    }
    if (lane == 0) {
      ld.len[cur ^ 1][g] = outlen < maxL ? outlen : maxL;
      if (outlen > maxL) atomicOr(err, kErrList);
    }
    __syncthreads();
    // clear the bits this grain set (all of them after an overflow, when some
    // set bits have no stored entry)
    if (outlen > maxL) {
      for (int w = lane; w < words; w += 64) bitmap[w] = 0u;
    } else {
      for (int i = lane; i < outlen; i += 64) {
        const int c = dst[i];
        atomicAnd(&bitmap[c >> 5], ~(1u << (c & 31)));
      }
    }
    __syncthreads();
    if (lane == 0) {
      bitmap[g >> 5] &= ~(1u << (g & 31));
      bitmap[0] &= ~1u;
    }
    __syncthreads();
  }
}

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

struct ActNb {
  int32_t id;
  int16_t x0, x1, y0, y1;
  int32_t nn;
  int64_t base;
};

template <class T>
__global__ __launch_bounds__(256) void k_comparison(SimParams sp, PlanDev P, const int32_t* __restrict__ state,
                                                    DescDev din, DescDev dout, ListDev ld, IdlDev ilo, IdlDev iln,
                                                    int32_t* newb) {
  __shared__ ActNb act[kMaxAct];
  __shared__ int s_wave[4];
  __shared__ int s_red[6][4];
  const int total = plan_total(P, kQLayers);
  const T* ain = (const T*)din.arena;
  T* aout = (T*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const T lim = (T)(0.7 * sp.delta);
  const int maxL = ld.max_list;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int c, g, kl;
    plan_tile_any(P, kQLayers, tile, c, g, kl);
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4], z1 = b[5];
    const int n = x1 - x0;
    const int k = z0 + kl;
    const int lsel = ld.sel[g] ^ 1;  // m_comparisonList
    const int32_t* lst = ld.buf[lsel] + (int64_t)g * maxL;
    const int len = ld.len[lsel][g];
    // compact the neighbours covering this layer, keeping list order
    int cnt = 0;
    for (int base = 0; base < len; base += 256) {
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
      cnt += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
      __syncthreads();
    }
    if (cnt > kMaxAct) cnt = kMaxAct;  // list capacity is below kMaxAct; guarded anyway
    // BoundaryIntersection on the output box (equal to the input box here)
    const int32_t* ob = dout.bounds + 6 * g;
    const int lo = sp.gb + sp.tube, hi = sp.ngp - sp.gb - sp.tube;
    const bool wall = !(ob[0] > lo && ob[3] < hi && ob[1] > lo && ob[4] < hi && ob[2] > lo && ob[5] < hi);
    const int32_t* olb = ilo.bounds + 6 * g;
    const int64_t oldsize = (int64_t)(olb[3] - olb[0]) * (olb[4] - olb[1]) * (olb[5] - olb[2]);
    const uint32_t* idl_old = ilo.arena + ilo.off[g];
    uint32_t* idl_new = iln.arena + iln.off[g];
    const T* own = ain + din.off[g];
    T* dst = aout + dout.off[g];
    // maxima start at -INT_MAX so that their negation (reduced with min below) cannot overflow
    int bx0 = INT_MAX, by0 = INT_MAX, bz0 = INT_MAX, bx1 = -INT_MAX, by1 = -INT_MAX, bz1 = -INT_MAX;
    const int64_t lbase = (int64_t)kl * n * n;
    for (int v = tid; v < n * n; v += 256) {
      const int i = y0 + v / n, j = x0 + v % n;
      T best = (T)(-1.0);
      uint32_t bid = 0;
      bool cov = false;
      for (int a = 0; a < cnt; ++a) {
        const ActNb& e = act[a];
        if (j >= e.x0 && j < e.x1 && i >= e.y0 && i < e.y1) {
          const T val = ain[e.base + (int64_t)(i - e.y0) * e.nn + (j - e.x0)];
          if (val > best) {
            best = val;
            bid = (uint32_t)e.id;
            cov = true;
          }
        }
      }
      if (wall) {
        const double dist = bc_dist(j, i, k, x0, y0, z0, x1, y1, sp.gb, sp.ngp);
        if (dist > -sp.gb) {
          const T dh = (T)(dist * sp.h);
          if (dh > best) {
            best = dh;
            bid = 0;
            cov = true;
          }
        }
      }
      const int64_t li = lbase + v;
      uint32_t idv;
      if (cov) idv = bid;
      else idv = li < oldsize ? idl_old[li] : 0u;
      idl_new[li] = idv;
      const T a_in = own[li];
      const T o = (a_in < T(0) ? -a_in : a_in) < lim ? (T)0.5 * (a_in - best) : a_in;
      dst[li] = o;
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
      int r[6];
      for (int q = 0; q < 6; ++q) r[q] = min(min(s_red[q][0], s_red[q][1]), min(s_red[q][2], s_red[q][3]));
      if (r[0] != INT_MAX) {
        int32_t* nb = newb + 6 * g;
        atomicMin(&nb[0], r[0]);
        atomicMin(&nb[1], r[1]);
        atomicMin(&nb[2], r[2]);
        atomicMax(&nb[3], -r[3]);
        atomicMax(&nb[4], -r[4]);
        atomicMax(&nb[5], -r[5]);
      }
    }
    __syncthreads();
  }
}

}  // namespace

void launch_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                  const DescDev& dout, const ListDev& ld, int32_t* newb, int32_t* err) {
  size_t lds = (size_t)((sp.N + 1 + 31) / 32) * sizeof(uint32_t);
  int grid = sp.N < 4096 ? sp.N : 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_lists, dim3(grid), dim3(64), lds, L.stream, sp, state, din, dout, ld, newb, err);
}

void launch_comparison(const Launch& L, const SimParams& sp, const PlanDev& P, const int32_t* state,
                       const DescDev& din, const DescDev& dout, const ListDev& ld, const IdlDev& il_old,
                       const IdlDev& il_new, int32_t* newb, bool fp32) {
  if (fp32)
    hipLaunchKernelGGL(k_comparison<float>, dim3(L.grid), dim3(256), 0, L.stream, sp, P, state, din, dout, ld,
                       il_old, il_new, newb);
  else
    hipLaunchKernelGGL(k_comparison<double>, dim3(L.grid), dim3(256), 0, L.stream, sp, P, state, din, dout, ld,
                       il_old, il_new, newb);
}

}  // namespace gg
