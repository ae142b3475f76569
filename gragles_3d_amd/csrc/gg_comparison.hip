// Comparison step of every grain: executeComparison (src/box.cpp:891-966) +
// BoundaryIntersection (:968-990) + boundaryCondition (:997-1088) +
// executeSetComparison (:814-878), fused per (grain, z-layer) tile, one launch
// per box-size class.
//
// Per tile the neighbours' values on the rectangles they share with the
// grain's layer are copied into LDS with independent, coalesced loads, then
// every lane takes the maximum over the neighbours for its voxels in
// comparison-list order with a strict '>' (ties keep the earlier neighbour),
// which is the voxel-wise result of the reference's neighbour-outer loop.
// Then the wall distance, IDLocal (new id, or the previous step's id at the
// same linear index when nothing covers the voxel), 0.5*(in - max) inside
// 0.7*delta, and the bounding box of phi >= 0.
#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kCT = 256;
constexpr int kWaves = kCT / 64;
constexpr int kLC = 128;            // comparison-list entries compacted per pass
constexpr int kStageBytes = 24576;  // LDS bytes of staged neighbour values per workgroup

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

// One tile = one z-layer of one grain.
//
// Active entries.  The comparison list is taken in chunks of kLC entries; the
// entries of a chunk whose box covers the layer and overlaps the grain's box in
// x and y are compacted in list order with the rectangle they share with the
// grain's layer, the rectangle's first arena index, and per wavefront a mask of
// the rectangle's 8x8 voxel blocks that wavefront owns.
//
// Staging.  The rectangles are laid end to end (a_pre = running element count)
// and copied into LDS in batches of BUF elements, every lane issuing SB
// independent loads at a time; a lane finds the entry of its element by one
// binary search per batch and then walks forward.
//
// Maximum.  The layer is cut into 8x8 blocks, block b owned by wavefront b % 4,
// one voxel per lane and block, so a lane's voxels stay in its registers for
// the whole tile.  For each entry of the batch, in list order, a wavefront
// visits only the blocks the entry's rectangle touches (one scalar bit test per
// block) and updates with a strict '>'.  A rectangle may straddle two batches:
// a voxel's staged index grows with the entry's list position, so each voxel
// still sees its neighbours in list order.
template <class T, int NMAX>
__global__ __launch_bounds__(kCT) void k_comparison(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                    ListDev ld, IdlDev ilo, IdlDev iln, int32_t* newb) {
  constexpr int NB = (NMAX + 7) / 8;                  // blocks per side, at most
  constexpr int VPT = (NB * NB + kWaves - 1) / kWaves;  // blocks per wavefront = voxels per lane
  static_assert(VPT <= 64, "a wavefront's block mask is one 64-bit word");
  constexpr int BUF = kStageBytes / (int)sizeof(T);
  constexpr int SB = sizeof(T) == 4 ? 16 : 8;
  __shared__ T sbuf[BUF];
  __shared__ int64_t a_base[kLC];  // arena index of the rectangle's first element
  __shared__ uint64_t a_wm[kWaves][kLC];  // blocks of the rectangle owned by each wavefront
  __shared__ int32_t a_id[kLC], a_nn[kLC], a_pre[kLC + 1];
  __shared__ uint32_t a_rect[kLC];   // local rectangle x0 | x1 << 8 | y0 << 16 | y1 << 24
  __shared__ uint32_t a_magic[kLC];  // ceil(2^32 / width) for the division by the width
  __shared__ int s_cnt[kLC / 64], s_area[kLC / 64];
  __shared__ int s_red[6][kWaves];
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const T lim = (T)(0.7 * sp.delta);
  const int maxL = ld.max_list;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4];
    const int n = x1 - x0, nn2 = n * n;
    const int nbx = (n + 7) >> 3;
    const int k = z0 + kl;
    const int lsel = ld.sel[g] ^ 1;  // m_comparisonList
    const int32_t* lst = ld.buf[lsel] + (int64_t)g * maxL;
    const int len = ld.len[lsel][g];

    // The lane's voxels, packed (i - y0) << 16 | (j - x0); -1 outside the layer.
    int ij[VPT];
    T best[VPT];
    uint32_t bid[VPT];
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      const int bb = wv + kWaves * r;
      const int by = bb / nbx, bx = bb - by * nbx;
      const int il = by * 8 + (lane >> 3), jl = bx * 8 + (lane & 7);
      ij[r] = (bb < nbx * nbx && il < n && jl < n) ? (il << 16) | jl : -1;
      best[r] = (T)(-1.0);
      bid[r] = 0xFFFFFFFFu;  // not covered
    }

    for (int lbase = 0; lbase < len; lbase += kLC) {
      // compact the active entries of this chunk, keeping list order
      const int idx = lbase + tid;
      bool keep = false;
      int area = 0, nb = 0, nn = 0, rx0 = 0, rx1 = 0, ry0 = 0, ry1 = 0;
      int64_t gbase = 0;
      if (tid < kLC && idx < len) {
        nb = lst[idx];
        const int32_t* q = din.bounds + 6 * nb;
        const int q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5];
        const int cx0 = max(x0, q0), cx1 = min(x1, q3), cy0 = max(y0, q1), cy1 = min(y1, q4);
        if (q2 <= k && k < q5 && cx0 < cx1 && cy0 < cy1) {
          keep = true;
          nn = q3 - q0;
          rx0 = cx0 - x0; rx1 = cx1 - x0; ry0 = cy0 - y0; ry1 = cy1 - y0;
          area = (rx1 - rx0) * (ry1 - ry0);
          gbase = din.off[nb] + ((int64_t)(k - q2) * nn + (cy0 - q1)) * nn + (cx0 - q0);
        }
      }
      const uint64_t m = __ballot(keep);
      int sc = area;  // inclusive scan of the areas over the wavefront
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(sc, off);
        if (lane >= off) sc += y;
      }
      if (lane == 63 && wv < kLC / 64) {
        s_cnt[wv] = __popcll(m);
        s_area[wv] = sc;
      }
      __syncthreads();
      if (keep) {
        int pos = __popcll(m & ((1ull << lane) - 1ull)), ab = sc - area;
        for (int w = 0; w < wv; ++w) {
          pos += s_cnt[w];
          ab += s_area[w];
        }
        uint64_t wm[kWaves] = {};
        for (int by = ry0 >> 3; by <= (ry1 - 1) >> 3; ++by)
          for (int bx = rx0 >> 3; bx <= (rx1 - 1) >> 3; ++bx) {
            const int bb = by * nbx + bx;
            wm[bb % kWaves] |= 1ull << (bb / kWaves);
          }
        for (int w = 0; w < kWaves; ++w) a_wm[w][pos] = wm[w];
        const int rw = rx1 - rx0;
        a_id[pos] = nb;
        a_nn[pos] = nn;
        a_rect[pos] = (uint32_t)rx0 | (uint32_t)rx1 << 8 | (uint32_t)ry0 << 16 | (uint32_t)ry1 << 24;
        a_magic[pos] = rw > 1 ? (uint32_t)(0xFFFFFFFFu / (uint32_t)rw + 1u) : 0u;
        a_base[pos] = gbase;
        a_pre[pos] = ab;
      }
      int cnt = 0, E = 0;
      for (int w = 0; w < kLC / 64; ++w) {
        cnt += s_cnt[w];
        E += s_area[w];
      }
      if (tid == 0) a_pre[cnt] = E;
      __syncthreads();

      int a_lo = 0;
      for (int s = 0; s < E; s += BUF) {
        const int e = min(s + BUF, E);
        int a = 0;  // entry of this lane's first element: last a with a_pre[a] <= s + tid
        {
          const int t = min(s + tid, e - 1);
#pragma unroll
          for (int st = kLC / 2; st > 0; st >>= 1)
            if (a + st < cnt && a_pre[a + st] <= t) a += st;
        }
        for (int t0 = s; t0 < e; t0 += kCT * SB) {
          T v[SB];
#pragma unroll
          for (int u = 0; u < SB; ++u) {
            const int t = t0 + u * kCT + tid;
            if (t < e) {
              while (a_pre[a + 1] <= t) ++a;  // a_pre[cnt] = E > t ends the walk
              const int r = t - a_pre[a];
              const uint32_t rc = a_rect[a];
              const int rw = (int)((rc >> 8) & 0xFFu) - (int)(rc & 0xFFu);
              const int qr = rw > 1 ? (int)__umulhi((uint32_t)r, a_magic[a]) : r;  // r / rw
              v[u] = ain[a_base[a] + (int64_t)qr * a_nn[a] + (r - qr * rw)];
            }
          }
#pragma unroll
          for (int u = 0; u < SB; ++u) {
            const int t = t0 + u * kCT + tid;
            if (t < e) sbuf[t - s] = v[u];
          }
        }
        __syncthreads();
        int aa = a_lo;
        for (; aa < cnt && a_pre[aa] < e; ++aa) {
          const uint64_t wmask = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a_wm[wv][aa] >> 32)) << 32) |
                                 (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a_wm[wv][aa]);
          const uint32_t rc = a_rect[aa];
          const int rx0 = rc & 0xFF, ry0 = (rc >> 16) & 0xFF;
          const int rw = (int)((rc >> 8) & 0xFFu) - rx0, rh = (int)(rc >> 24) - ry0;
          const int pre = a_pre[aa] - s;  // sbuf index of the rectangle's first element (may be < 0)
          const uint32_t lim_e = (uint32_t)(e - s);
          const uint32_t id = (uint32_t)a_id[aa];
#pragma unroll
          for (int r = 0; r < VPT; ++r) {
            if ((wmask >> r) & 1ull) {
              const int dx = (ij[r] & 0xFFFF) - rx0, dy = (ij[r] >> 16) - ry0;
              const int el = pre + dy * rw + dx;
              if ((uint32_t)dx < (uint32_t)rw && (uint32_t)dy < (uint32_t)rh && (uint32_t)el < lim_e) {
                const T val = sbuf[el];
                if (val > best[r]) {
                  best[r] = val;
                  bid[r] = id;
                }
              }
            }
          }
        }
        // the next batch starts with the entry that straddles e, if any
        a_lo = a_pre[aa] > e ? aa - 1 : aa;
        __syncthreads();
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
        for (int w = 1; w < kWaves; ++w) rr[q] = min(rr[q], s_red[q][w]);
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
void comparison_class(const Launch& L, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                      const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  switch (cls) {
    case 0:
      hipLaunchKernelGGL((k_comparison<T, 32>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 0, din, dout, ld, ilo,
                         iln, newb);
      break;
    case 1:
      hipLaunchKernelGGL((k_comparison<T, 48>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 1, din, dout, ld, ilo,
                         iln, newb);
      break;
    case 2:
      hipLaunchKernelGGL((k_comparison<T, 64>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 2, din, dout, ld, ilo,
                         iln, newb);
      break;
    default:
      hipLaunchKernelGGL((k_comparison<T, 96>), dim3(L.grid), dim3(kCT), 0, L.stream, sp, P, 3, din, dout, ld, ilo,
                         iln, newb);
  }
}

}  // namespace

// One size class per launch (boxes n <= 32, 48, 64, 96 for cls 0..3).
void launch_comparison(const Launch& L, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                       const DescDev& dout, const ListDev& ld, const IdlDev& il_old, const IdlDev& il_new,
                       int32_t* newb, bool fp32) {
  if (fp32) comparison_class<float>(L, cls, sp, P, din, dout, ld, il_old, il_new, newb);
  else comparison_class<double>(L, cls, sp, P, din, dout, ld, il_old, il_new, newb);
}

}  // namespace gg
