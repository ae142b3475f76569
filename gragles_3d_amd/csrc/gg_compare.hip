// Comparison lists of every grain.
//
// k_lists: LSbox::computeSecondOrderNeighbours (src/box.cpp:1090-1134) with
// checkIntersection (src/box.cpp:880-889).  One wavefront per grain; a bitmap
// of grain ids in LDS replaces the reference's linear duplicate scan, and the
// candidates are appended strictly in the reference's order (previous list
// entries that still exist, then neighbours-of-neighbours in traversal order,
// first occurrence wins), because that order decides comparison ties.

#include "gg_internal.h"

namespace gg {
namespace {

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
        // A list holds no id twice, so the ids of one chunk are distinct and
        // "first occurrence" reduces to "not yet in the bitmap"; the ballot
        // prefix appends the fresh ones in lane (= traversal) order.
        const bool fresh = p && ((bitmap[cid >> 5] >> (cid & 31)) & 1u) == 0u;
        const uint64_t m = __ballot(fresh);
        const int pos = outlen + __popcll(m & ((1ull << lane) - 1ull));
        if (fresh) {
          if (pos < maxL) dst[pos] = cid;
          atomicOr(&bitmap[cid >> 5], 1u << (cid & 31));
        }
        outlen += __popcll(m);
        __syncthreads();
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

}  // namespace

void launch_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                  const DescDev& dout, const ListDev& ld, int32_t* newb, int32_t* err) {
  size_t lds = (size_t)((sp.N + 1 + 31) / 32) * sizeof(uint32_t);
  int grid = sp.N < 4096 ? sp.N : 4096;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_lists, dim3(grid), dim3(64), lds, L.stream, sp, state, din, dout, ld, newb, err);
}

}  // namespace gg
