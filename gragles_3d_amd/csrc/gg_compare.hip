// Comparison lists of every grain.
//
// k_lists: LSbox::computeSecondOrderNeighbours (src/box.cpp:1090-1134) with
// checkIntersection (src/box.cpp:880-889).  One workgroup per grain.  The
// result is the reference's sequence: previous list entries that still exist,
// in order, then the neighbours-of-neighbours whose input boxes intersect, in
// traversal order, each id at its first occurrence (that order decides
// comparison ties).  Instead of walking the candidates one after another, the
// workgroup checks them all at once: the neighbours' lists form one flat
// sequence of candidates (a prefix of their lengths maps a position to its list
// and entry), every valid candidate enters an LDS hash table keyed by grain id
// that keeps the smallest position of each id (previous entries hold position
// -1), and a candidate is appended iff its position is the one the table kept;
// a workgroup scan over the positions puts the appended ids in order.  The
// candidates are taken in windows of kWin positions; LDS use depends on the list
// capacity only, not on the number of grains.

#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kLT = 256;            // threads per grain
constexpr int kHBits = 11;          // hash table of 2048 ids (at most 1024 are kept)
constexpr int kH = 1 << kHBits;
constexpr int kWin = 2048;          // candidate positions per window
constexpr int kEmpty = -1;

// Workgroup exclusive scan of v (kLT threads); *total gets the sum.  Two barriers.
__device__ inline int block_scan(int v, int* s_w, int& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int sc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(sc, off);
    if (lane >= off) sc += y;
  }
  if (lane == 63) s_w[wv] = sc;
  __syncthreads();
  int pre = 0;
  total = 0;
  for (int w = 0; w < kLT / 64; ++w) {
    if (w < wv) pre += s_w[w];
    total += s_w[w];
  }
  __syncthreads();
  return pre + sc - v;
}

__device__ inline uint32_t hslot(int id) { return ((uint32_t)id * 2654435761u) >> (32 - kHBits); }

// Enter id with position t (keeping the smallest); false when the table is full.
__device__ inline bool hash_put(int* keys, int* pos, int id, int t) {
  uint32_t h = hslot(id);
  for (int probe = 0; probe < kH; ++probe) {
    const int old = atomicCAS(&keys[h], kEmpty, id);
    if (old == kEmpty || old == id) {
      atomicMin(&pos[h], t);
      return true;
    }
    h = (h + 1) & (kH - 1);
  }
  return false;
}

// Smallest position entered for id (id must be in the table).
__device__ inline int hash_pos(const int* keys, const int* pos, int id) {
  uint32_t h = hslot(id);
  for (int probe = 0; probe < kH; ++probe) {
    if (keys[h] == id) return pos[h];
    h = (h + 1) & (kH - 1);
  }
  return INT_MIN;
}

__global__ __launch_bounds__(kLT) void k_lists(SimParams sp, const int32_t* __restrict__ state, DescDev din,
                                               DescDev dout, ListDev ld, int32_t* newb, int32_t* err) {
  extern __shared__ int32_t s_pre[];  // [max_list + 1] start of each neighbour's list in the flat sequence
  __shared__ int keys[kH], pos[kH];
  __shared__ int s_cid[kWin];
  __shared__ int s_w[kLT / 64];
  const int maxL = ld.max_list;
  const int N = sp.N;
  const int tid = threadIdx.x;
  for (int g = 1 + blockIdx.x; g <= N; g += gridDim.x) {
    if (state[g] != kAlive) continue;
    if (tid == 0) {
      // executeSetComparison's starting values m_newXMin = out.maxX ... (src/box.cpp:815-820)
      const int32_t* ob = dout.bounds + 6 * g;
      int32_t* nb = newb + 6 * g;
      nb[0] = ob[3]; nb[1] = ob[4]; nb[2] = ob[5];
      nb[3] = ob[0]; nb[4] = ob[1]; nb[5] = ob[2];
    }
    for (int i = tid; i < kH; i += kLT) {
      keys[i] = kEmpty;
      pos[i] = INT_MAX;
    }
    __syncthreads();
    const int cur = ld.sel[g];
    const int32_t* src = ld.buf[cur] + (int64_t)g * maxL;
    const int srclen = ld.len[cur][g];
    int32_t* dst = ld.buf[cur ^ 1] + (int64_t)g * maxL;
    const int32_t* gbnd = din.bounds + 6 * g;
    const int gx0 = gbnd[0], gy0 = gbnd[1], gz0 = gbnd[2], gx1 = gbnd[3], gy1 = gbnd[4], gz1 = gbnd[5];
    bool full = false;
    int outlen = 0, run = 0;
    // previous neighbours that still exist, in order (a list holds no id twice),
    // and the flat positions of the neighbours' lists
    for (int base = 0; base < srclen; base += kLT) {
      const int idx = base + tid;
      int id = 0, nl = 0;
      bool keep = false;
      if (idx < srclen) {
        id = src[idx];
        keep = state[id] == kAlive;
        nl = ld.len[ld.sel[id]][id];
      }
      int cnt, tot;
      const int rank = block_scan(keep ? 1 : 0, s_w, cnt);
      const int start = block_scan(nl, s_w, tot);
      if (keep) {
        if (outlen + rank < maxL) dst[outlen + rank] = id;
        full |= !hash_put(keys, pos, id, -1);
      }
      if (idx < srclen) s_pre[idx] = run + start;
      outlen += cnt;
      run += tot;
    }
    if (tid == 0) s_pre[srclen] = run;
    __syncthreads();
    // neighbours of neighbours whose input boxes intersect (inclusive bounds)
    for (int w0 = 0; w0 < run; w0 += kWin) {
      const int wn = min(kWin, run - w0);
      for (int i = tid; i < wn; i += kLT) {
        const int t = w0 + i;
        int q = 0;  // last list with s_pre[q] <= t (empty lists share a start)
        for (int st = 512; st > 0; st >>= 1)
          if (q + st < srclen && s_pre[q + st] <= t) q += st;
        const int nbid = src[q];
        const int cid = ld.buf[ld.sel[nbid]][(int64_t)nbid * maxL + (t - s_pre[q])];
        bool p = false;
        // the grain itself and the wall never become candidates (src/box.cpp:1118-1122)
        if (cid != g && cid != 0 && state[cid] == kAlive) {
          const int32_t* c6 = din.bounds + 6 * cid;
          p = !(gx0 > c6[3] || gx1 < c6[0] || gy0 > c6[4] || gy1 < c6[1] || gz0 > c6[5] || gz1 < c6[2]);
        }
        s_cid[i] = p ? cid : kEmpty;
        if (p) full |= !hash_put(keys, pos, cid, t);
      }
      __syncthreads();
      // in window order: a candidate is new where its position is the smallest of its id
      for (int base = 0; base < wn; base += kLT) {
        const int i = base + tid;
        bool first = false;
        int cid = kEmpty;
        if (i < wn) {
          cid = s_cid[i];
          first = cid != kEmpty && hash_pos(keys, pos, cid) == w0 + i;
        }
        int cnt;
        const int rank = block_scan(first ? 1 : 0, s_w, cnt);
        if (first && outlen + rank < maxL) dst[outlen + rank] = cid;
        outlen += cnt;
      }
      __syncthreads();
    }
    if (full) atomicOr(err, kErrList);
    if (tid == 0) {
      ld.len[cur ^ 1][g] = outlen < maxL ? outlen : maxL;
      if (outlen > maxL) atomicOr(err, kErrList);
    }
    __syncthreads();
  }
}

}  // namespace

// Initial m_secondOrderNeighbours from the boxes alone (grainhdl::find_neighbors,
// src/grainhdl.cpp:1029-1054, with LSbox::computeDirectNeighbours,
// src/box.cpp:2871-2887): every other alive grain whose cubed box overlaps the
// grain's with inclusive bounds, in ascending id (the reference's R-tree
// traversal order is not reproduced; the host routine gg_initial_neighbour_lists
// uses the same order).  One wavefront per grain, all grains tested 64 at a
// time: O(N^2) tests, a one-off at set-up.
__global__ __launch_bounds__(64) void k_init_lists(SimParams sp, const int32_t* __restrict__ state, DescDev d,
                                                   ListDev ld, int32_t* err) {
  const int N = sp.N, maxL = ld.max_list, lane = threadIdx.x;
  for (int g = 1 + blockIdx.x; g <= N; g += gridDim.x) {
    int out = 0;
    if (state[g] == kAlive) {
      const int32_t* a = d.bounds + 6 * g;
      const int a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5];
      int32_t* dst = ld.buf[0] + (int64_t)g * maxL;
      for (int base = 1; base <= N; base += 64) {
        const int h = base + lane;
        bool p = false;
        if (h <= N && h != g && state[h] == kAlive) {
          const int32_t* b = d.bounds + 6 * h;
          p = !(a0 > b[3] || a3 < b[0] || a1 > b[4] || a4 < b[1] || a2 > b[5] || a5 < b[2]);
        }
        const uint64_t m = __ballot(p);
        const int pos = out + __popcll(m & ((1ull << lane) - 1ull));
        if (p && pos < maxL) dst[pos] = h;
        out += __popcll(m);
      }
    }
    if (lane == 0) {
      ld.len[0][g] = out < maxL ? out : maxL;
      ld.sel[g] = 0;
      if (out > maxL) atomicOr(err, kErrList);
    }
  }
}

void launch_init_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& d,
                       const ListDev& ld, int32_t* err) {
  int grid = sp.N < 65536 ? sp.N : 65536;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_init_lists, dim3(grid), dim3(64), 0, L.stream, sp, state, d, ld, err);
}

void launch_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                  const DescDev& dout, const ListDev& ld, int32_t* newb, int32_t* err) {
  size_t lds = (size_t)(ld.max_list + 1) * sizeof(int32_t);
  int grid = sp.N < 65536 ? sp.N : 65536;  // one workgroup per grain
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(k_lists, dim3(grid), dim3(kLT), lds, L.stream, sp, state, din, dout, ld, newb, err);
}

}  // namespace gg
