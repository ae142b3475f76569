// Comparison step of every grain: executeComparison (src/box.cpp:891-966) +
// BoundaryIntersection (:968-990) + boundaryCondition (:997-1088) +
// executeSetComparison (:814-878), fused, one launch per box-size class.
//
// The work unit is a wavefront tile: 64 consecutive columns (y, x) of a
// grain's box over a run of z-layers (the run length is tuned per size class).  Each voxel takes the maximum
// over the grain's neighbours in comparison-list order with a strict '>'
// (ties keep the earlier neighbour), which is the voxel-wise result of the
// reference's neighbour-outer loop; then the wall distance, IDLocal (new id,
// or the previous step's id at the same linear index when nothing covers the
// voxel), 0.5*(in - max) inside 0.7*delta, and the bounding box of phi >= 0.
#include "gg_internal.h"

namespace gg {
namespace {
// Value of boundaryCondition's `dist` at voxel (j,i,k): the reference assigns
// it in every branch except when all three axis distances are >= 0 and one is
// 0; there it keeps the value of the previous voxel in its (k,i,j) loop over
// the box (starting from 0).  Walk back in that order to the last assignment.
__device__ double bc_dist(int j, int i, int k, int x0, int y0, int z0, int x1, int y1, int gb,
                                                  int m) {
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

// True when boundaryCondition leaves voxel (j,i,k) alone: all three axis
// distances are negative (dist is their maximum, assigned in this voxel) and
// dist <= -gb.  Integer form of the first lines of bc_dist; most voxels of a
// box that touches the domain wall take this exit.
__device__ inline bool deep_inside(int j, int i, int k, int gb, int m) {
  const int dXm = gb - 1 - j, dXM = j - (m - 1 - gb);
  const int dYm = gb - 1 - i, dYM = i - (m - 1 - gb);
  const int dZm = gb - 1 - k, dZM = k - (m - 1 - gb);
  const int dX = abs(dXm) < abs(dXM) ? dXm : dXM;
  const int dY = abs(dYm) < abs(dYM) ? dYm : dYM;
  const int dZ = abs(dZm) < abs(dZM) ? dZm : dZM;
  return dX < 0 && dY < 0 && dZ < 0 && max(dX, max(dY, dZ)) <= -gb;
}

// One wavefront per tile (grain, 64 consecutive columns (y, x) of its box,
// run of ZL z-layers), tiles in the plan's order: grains in spatial order, so
// the wavefronts that run together read the same neighbour boxes.  Lane l owns column cb*64 + l (row-major in the
// layer, so loads and stores are contiguous across lanes) and keeps the
// running maximum and its id for each of the run's layers in registers.  The
// neighbour table is walked once per wavefront, in list order: an entry whose
// box overlaps the run in z and whose rectangle covers some lane's column
// gives each covered lane its offset in the neighbour's layer once, then its
// values over the run's layers are gathered QB at a time (independent loads,
// the layer index clamped into the entry's z range so every address is inside
// the neighbour's box; lanes off the rectangle read its first voxel) and the
// maxima are updated with a strict '>' (ties keep the earlier entry: the
// reference's order).  The per-entry set-up is paid once per ZL layers, there
// is no LDS and no barrier.  Then, per voxel, the wall distance, IDLocal (new
// id or the stale one), the set-comparison and the bounding box, as above.
// One wavefront tile: column block cb (64 consecutive columns (y, x) of grain
// g's box) over run zc of ZL z-layers.
// B16: the winning ids of the run's layers are kept as 16-bit halves, two per
// register (grain ids below 0xFFFF; 0xFFFF stands for "not covered").  With
// the spatial kernel's request of eight wavefronts per SIMD (64 vector
// registers) the FP32 body spills about ten registers, outside the gather
// loop, and still runs faster than at seven wavefronts with unpacked ids
// (3.33 against 3.65 ms at 8000 grains, measured with a grid-stride launch of
// 64 workgroups per CU).
template <bool B16>
__device__ inline uint32_t bid_get(const uint32_t* bid, int q) {
  if constexpr (!B16) return bid[q];
  const uint32_t v = (bid[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
  return v == 0xFFFFu ? 0xFFFFFFFFu : v;
}
template <bool B16>
__device__ inline uint32_t bid_with(uint32_t word, int q, uint32_t id) {
  if constexpr (!B16) return id;
  return (q & 1) ? ((word & 0xFFFFu) | (id << 16)) : ((word & 0xFFFF0000u) | (id & 0xFFFFu));
}

template <class T, int ZL, int QB, bool B16 = false>
__device__ inline void cmp_tile(const SimParams& sp, int g, int cb, int zc, const DescDev& din, const DescDev& dout,
                                const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  static_assert(ZL % QB == 0, "a run is whole gather batches");
  const int lane = threadIdx.x & 63;
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const T lim = (T)(0.7 * sp.delta);
  const int maxL = ld.max_list;
  const int32_t* b = din.bounds + 6 * g;
  const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4];
  const int n = x1 - x0, nn2 = n * n;
  const int qlo = zc * ZL;
  if (qlo >= n) return;
  const int qn = min(ZL, n - qlo);  // layers of this run
  const int zlo = z0 + qlo;         // its first global layer
  const int col = cb * 64 + lane;
  const bool act = col < nn2;
  const int yl = act ? col / n : 0, xl = act ? col - yl * n : 0;
  T best[ZL];
  constexpr int NB = B16 ? (ZL + 1) / 2 : ZL;
  uint32_t bid[NB];
#pragma unroll
  for (int q = 0; q < ZL; ++q) {
    best[q] = (T)(-1.0);
    if (q < NB) bid[q] = 0xFFFFFFFFu;  // not covered
  }
  const int4* tab = ld.tab + (int64_t)g * maxL;
  const int32_t* tid_of = ld.tab_id + (int64_t)g * maxL;
  const int len = ld.tab_len[g];
  for (int pb = 0; pb < len; pb += 64) {
    const int idx = pb + lane;
    bool keep = false;
    int nb = 0, q2 = 0, q5 = 0;
    int4 e = make_int4(0, 0, 0, 0);
    if (idx < len) {
      e = tab[idx];
      nb = tid_of[idx];
      q2 = e.w & 0xFFFF;
      q5 = (int)((uint32_t)e.w >> 16);
      keep = q2 < zlo + qn && q5 > zlo;  // the neighbour's box meets the run's layers
    }
    uint64_t m = __ballot(keep);
    while (m) {
      const int a = (int)__builtin_ctzll(m);  // next entry in list order
      m &= m - 1;
      const uint32_t rc = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)e.z, a);
      const int rx0 = rc & 0xFF, rx1 = (rc >> 8) & 0xFF, ry0 = (rc >> 16) & 0xFF, ry1 = rc >> 24;
      const bool cov = act & ((uint32_t)(xl - rx0) < (uint32_t)(rx1 - rx0)) & ((uint32_t)(yl - ry0) < (uint32_t)(ry1 - ry0));
      const uint64_t cmask = __ballot(cov);
      if (cmask == 0) continue;  // the rectangle misses this wavefront's columns
      const uint32_t id = (uint32_t)__builtin_amdgcn_readlane(nb, a);
      const uint32_t elo = (uint32_t)__builtin_amdgcn_readlane(e.x, a);
      const uint32_t ehi = (uint32_t)__builtin_amdgcn_readlane(e.y, a);
      const int nn = (int)(ehi >> 16);
      const int64_t B = (int64_t)((uint64_t)(ehi & 0xFFFFu) << 32 | elo);  // rectangle's first voxel, layer q2
      const int eq2 = __builtin_amdgcn_readlane(q2, a), eq5 = __builtin_amdgcn_readlane(q5, a);
      const int qa = max(eq2, zlo) - zlo, qb = min(eq5, zlo + qn) - zlo;  // run layers the entry covers
      const int ls = nn * nn;
      // layer zlo + qa of the neighbour; element q of the run is at (q - qa) * ls + off from it
      const T* __restrict__ src = ain + B + (int64_t)(zlo + qa - eq2) * ls;
      // lanes off the rectangle read the first covered lane's voxel: a line the
      // gather fetches anyway (the rectangle's first voxel of every layer would
      // be one more line per gather)
      const uint32_t offc = (uint32_t)((yl - ry0) * nn + (xl - rx0));
      const uint32_t off0 = (uint32_t)__builtin_amdgcn_readlane((int)offc, (int)__builtin_ctzll(cmask));
      const uint32_t off = cov ? offc : off0;
#pragma unroll
      for (int c0 = 0; c0 < ZL; c0 += QB) {
        if (c0 >= qb || c0 + QB <= qa) continue;
        T v[QB];
#pragma unroll
        for (int u = 0; u < QB; ++u) {
          const int qc = min(max(c0 + u, qa), qb - 1) - qa;
          // 32-bit byte offset from a wavefront-uniform base
          v[u] = *(const T*)((const char*)src + (uint32_t)((qc * ls + (int)off) * (int)sizeof(T)));
        }
#pragma unroll
        for (int u = 0; u < QB; ++u) {
          const int q = c0 + u;
          const bool up = cov & (q >= qa) & (q < qb) & (v[u] > best[q]);
          best[q] = up ? v[u] : best[q];
          const int w = B16 ? (q >> 1) : q;
          bid[w] = up ? bid_with<B16>(bid[w], q, id) : bid[w];
        }
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
  const T* own = ain + din.off[g];
  T* dst = aout + dout.off[g];
  const int i = y0 + yl, j = x0 + xl;
  int kmin = INT_MAX, kmax = -INT_MAX;
#pragma unroll
  for (int q = 0; q < ZL; ++q) {
    if (q >= qn || !act) continue;
    const int k = zlo + q;
    if (wall && !deep_inside(j, i, k, sp.gb, sp.ngp)) {
      const double dist = bc_dist(j, i, k, x0, y0, z0, x1, y1, sp.gb, sp.ngp);
      if (dist > -sp.gb) {
        const T dh = (T)(dist * sp.h);
        if (dh > best[q]) {
          best[q] = dh;
          const int w = B16 ? (q >> 1) : q;
          bid[w] = bid_with<B16>(bid[w], q, 0u);
        }
      }
    }
    const int64_t li = (int64_t)(qlo + q) * nn2 + col;
    uint32_t idv = bid_get<B16>(bid, q);
    if (idv == 0xFFFFFFFFu) idv = li < oldsize ? idl_old[li] : 0u;
    idl_new[li] = idv;
    const T a_in = own[li];
    const T o = (a_in < T(0) ? -a_in : a_in) < lim ? (T)0.5 * (a_in - best[q]) : a_in;
    dst[li] = o;
    if (o >= T(0)) {
      kmin = min(kmin, k);
      kmax = max(kmax, k);
    }
  }
  // bounding box of the voxels >= 0: the lane's column spans (j, i) x [kmin, kmax]
  const bool any = kmin != INT_MAX;
  int vals[6] = {any ? j : INT_MAX, any ? i : INT_MAX, kmin, any ? -j : INT_MAX, any ? -i : INT_MAX, -kmax};
  for (int q = 0; q < 6; ++q) {
    int x = vals[q];
    for (int o = 32; o > 0; o >>= 1) x = min(x, __shfl_xor(x, o));
    vals[q] = x;
  }
  if (lane == 0 && vals[0] != INT_MAX) {
    int32_t* nbx = newb + 6 * g;
    atomicMin(&nbx[0], vals[0]);
    atomicMin(&nbx[1], vals[1]);
    atomicMin(&nbx[2], vals[2]);
    atomicMax(&nbx[3], -vals[3]);
    atomicMax(&nbx[4], -vals[4]);
    atomicMax(&nbx[5], -vals[5]);
  }
}

template <class T, int NMAX, int ZL, int QB, bool ZMAJ>
__global__ __launch_bounds__(256) void k_cmp_cols(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                  ListDev ld, IdlDev ilo, IdlDev iln, int32_t* newb) {
  constexpr int ZCH = (NMAX + ZL - 1) / ZL;  // layer runs per column block
  // wavefront tiles: the plan's column blocks (grains in spatial order), each
  // with its ZCH runs next to each other
  const int total = P.totals[kQCols * kClasses + cls] * ZCH;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: everything per tile is scalar
  const TileRange tr = xcd_tiles((total + 3) >> 2);  // workgroups take 4 consecutive wavefront tiles
  for (int bt = tr.first; bt < tr.end; bt += tr.step) {
    const int wt = bt * 4 + wv;
    if (wt >= total) continue;
    const int ct = wt / ZCH;
    int zc = wt - ct * ZCH;
    int g, cb;
    plan_tile(P, cls, kQCols, ct, g, cb);
    if constexpr (ZMAJ) {
      // the grain's wavefront tiles in run-major order: consecutive tiles (the
      // wavefronts of one workgroup, then the next workgroups) take consecutive
      // column blocks of one run of layers, so they read the same neighbour rows
      const int32_t* b = din.bounds + 6 * g;
      const int n = b[3] - b[0];
      const int nbk = (n * n + kColBlock - 1) / kColBlock;
      const int lt = wt - (ct - cb) * ZCH;  // tile index inside the grain
      zc = lt / nbk;
      cb = lt - zc * nbk;
    }
    cmp_tile<T, ZL, QB>(sp, g, cb, zc, din, dout, ld, ilo, iln, newb);
  }
}

// Neighbour table of every alive grain (layout in gg_internal.h): the entries
// of m_comparisonList whose input box overlaps the grain's box in x, y and z,
// in list order, with the shared x/y rectangle and where it starts in the
// arena.  One wavefront per grain; the comparison tiles of the grain's layers
// then read one table entry per neighbour instead of the list, the
// neighbour's bounds and its arena offset one after the other.
__global__ __launch_bounds__(64) void k_cmp_table(SimParams sp, const int32_t* __restrict__ state, DescDev din,
                                                  ListDev ld, int coarsen) {
  const int lane = threadIdx.x;
  const int maxL = ld.max_list;
  for (int g = 1 + blockIdx.x; g <= sp.N; g += gridDim.x) {
    if (state[g] != kAlive) {
      if (lane == 0) ld.tab_len[g] = 0;
      continue;
    }
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4], z1 = b[5];
    const int lsel = ld.sel[g] ^ (coarsen ? 0 : 1);  // m_comparisonList (coarsening: m_secondOrderNeighbours)
    const int32_t* lst = ld.buf[lsel] + (int64_t)g * maxL;
    const int len = ld.len[lsel][g];
    int4* tab = ld.tab + (int64_t)g * maxL;
    int32_t* tid_of = ld.tab_id + (int64_t)g * maxL;
    int out = 0;
    for (int base = 0; base < len; base += 64) {
      const int idx = base + lane;
      bool keep = false;
      int nb = 0;
      int4 e = make_int4(0, 0, 0, 0);
      if (idx < len && !(coarsen && state[lst[idx]] != kAlive)) {
        nb = lst[idx];
        const int32_t* q = din.bounds + 6 * nb;
        const int q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5];
        const int cx0 = max(x0, q0), cx1 = min(x1, q3), cy0 = max(y0, q1), cy1 = min(y1, q4);
        if (cx0 < cx1 && cy0 < cy1 && max(z0, q2) < min(z1, q5)) {
          keep = true;
          const int nn = q3 - q0;
          const int64_t B = din.off[nb] + (int64_t)(cy0 - q1) * nn + (cx0 - q0);
          e = make_int4((int)(uint32_t)B, (int)(((uint32_t)((uint64_t)B >> 32) & 0xFFFFu) | (uint32_t)nn << 16),
                        (int)((uint32_t)(cx0 - x0) | (uint32_t)(cx1 - x0) << 8 | (uint32_t)(cy0 - y0) << 16 |
                              (uint32_t)(cy1 - y0) << 24),
                        (int)((uint32_t)q2 | (uint32_t)q5 << 16));
        }
      }
      const uint64_t m = __ballot(keep);
      if (keep) {
        const int pos = out + __popcll(m & ((1ull << lane) - 1ull));
        tab[pos] = e;
        tid_of[pos] = nb;
      }
      out += __popcll(m);
    }
    if (lane == 0) ld.tab_len[g] = out;
  }
}

// Workgroups per CU of the spatially ordered launch: about one per 4 wavefront
// tiles at 8000 grains (the dispatcher then balances tiles of uneven cost);
// sweep of the comparison segment: 32: 3.59, 64: 3.35, 128: 3.27, 256: 3.18,
// 512: 3.16, 1024: 3.18 ms.
constexpr int kColGridPerCUSpace = 512;
// ---- spatial work list ----------------------------------------------------
// The wavefront tiles of all grains that read the same neighbour rows are put
// next to each other: a tile's key is the cell of its centre (x: the middle of
// the box, whose rows a 64-column block spans; y: its middle row; z: the middle
// of its run).  The launch then walks the keys in order, each XCD a contiguous
// range, so the ~7 grains whose boxes cover a voxel read its neighbour values
// while they are in that XCD's L2, instead of once per grain from HBM.
__device__ inline int space_key(const CmpSpaceDev& W, int x, int y, int z) {
  const int cx = x >> 4, cy = y >> 3, cz = z >> 3;
  const int brick = ((cz >> 3) * W.by + (cy >> 3)) * W.bx + (cx >> 3);
  int m = 0;
#pragma unroll
  for (int b = 0; b < 3; ++b)
    m |= ((cx >> b) & 1) << (3 * b) | ((cy >> b) & 1) << (3 * b + 1) | ((cz >> b) & 1) << (3 * b + 2);
  return brick * 512 + m;
}

__global__ __launch_bounds__(256) void k_space_clear(CmpSpaceDev W) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= W.nkeys; i += gridDim.x * blockDim.x) W.cnt[i] = 0;
}

// One wavefront per planned grain, lanes over its tiles: count per key, or
// (FILL, after the scan) take a slot in the key's bucket.  Slots inside a
// bucket come in atomic order; the tiles write disjoint voxels and the bounds
// atomics are minima and maxima, so the results do not depend on it.
template <bool FILL>
__global__ __launch_bounds__(256) void k_space_bin(SimParams sp, PlanDev P, DescDev din, CmpSpaceDev W, int zl,
                                                   int32_t* err) {
  const int lane = threadIdx.x & 63;
  const int ng = P.cls_start[kClasses - 1] + P.cls_count[kClasses - 1];
  const int lim = sp.ngp;
  for (int i = blockIdx.x * 4 + (threadIdx.x >> 6); i < ng; i += gridDim.x * 4) {
    const int g = P.grains[i];
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], n = b[3] - b[0], nn2 = n * n;
    const int nbk = (nn2 + kColBlock - 1) / kColBlock, nzc = (n + zl - 1) / zl;
    const int xc = min(max(x0 + (n >> 1), 0), lim);
    for (int t = lane; t < nbk * nzc; t += 64) {
      const int zc = t / nbk, cb = t - zc * nbk;
      const int c0 = cb * kColBlock, c1 = min(nn2, c0 + kColBlock) - 1;
      const int yc = min(max(y0 + ((c0 / n + c1 / n) >> 1), 0), lim);
      const int zq = min(zl, n - zc * zl);
      const int zm = min(max(z0 + zc * zl + (zq >> 1), 0), lim);
      const int key = space_key(W, xc, yc, zm);
      if constexpr (!FILL) {
        atomicAdd(&W.cnt[key], 1);
      } else {
        const int pos = atomicAdd(&W.cnt[key], 1);
        if (pos < W.cap) W.wl[pos] = make_int2(g, cb | zc << 16);
        else atomicOr(err, kErrCmpWork);
      }
    }
  }
}

// Exclusive scan in place over a[0..n), a[n] = total: blocks of kScanBlock
// elements (4 contiguous per thread), then the block sums, then their offsets.
constexpr int kScanBlock = 4096;
constexpr int kScanT = kScanBlock / 4;

__device__ inline int block_excl_scan(int v, int* s_w, int& total) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o);
    if (lane >= o) inc += y;
  }
  if (lane == 63) s_w[wv] = inc;
  __syncthreads();
  int base = 0;
  total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    if (w < wv) base += s_w[w];
    total += s_w[w];
  }
  __syncthreads();  // s_w is reused by the next call
  return base + inc - v;
}

__global__ __launch_bounds__(kScanT) void k_scan_blocks(int32_t* a, int n, int32_t* bsum) {
  __shared__ int s_w[kScanT / 64];
  const int i0 = blockIdx.x * kScanBlock + threadIdx.x * 4;
  int v[4], sum = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    v[u] = i0 + u < n ? a[i0 + u] : 0;
    sum += v[u];
  }
  int tot;
  int run = block_excl_scan(sum, s_w, tot);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    if (i0 + u < n) a[i0 + u] = run;
    run += v[u];
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanT) void k_scan_sums(int32_t* a, int n, int32_t* bsum, int nb) {
  __shared__ int s_w[kScanT / 64];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kScanT) {
    const int b = b0 + (int)threadIdx.x;
    const int v = b < nb ? bsum[b] : 0;
    int tot;
    const int ex = block_excl_scan(v, s_w, tot);
    if (b < nb) bsum[b] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) a[n] = carry;
}

__global__ __launch_bounds__(kScanT) void k_scan_add(int32_t* a, int n, const int32_t* __restrict__ bsum) {
  const int add = bsum[blockIdx.x];
  const int i0 = blockIdx.x * kScanBlock + threadIdx.x * 4;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (i0 + u < n) a[i0 + u] += add;
}

template <class T, int ZL, int QB, bool B16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(B16 && sizeof(T) == 4 ? 8 : 1))) void k_cmp_space(SimParams sp, CmpSpaceDev W, DescDev din, DescDev dout,
                                                   ListDev ld, IdlDev ilo, IdlDev iln, int32_t* newb) {
  const int total = min(W.cnt[W.nkeys], W.cap);
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const TileRange tr = xcd_tiles((total + 3) >> 2);
  for (int bt = tr.first; bt < tr.end; bt += tr.step) {
    const int wt = bt * 4 + wv;
    if (wt >= total) continue;
    const int2 e = W.wl[wt];
    cmp_tile<T, ZL, QB, B16>(sp, e.x, e.y & 0xFFFF, e.y >> 16, din, dout, ld, ilo, iln, newb);
  }
}

constexpr int kSpaceShapes[kCmpSpaceShapes][2] = {{8, 8}, {16, 8}, {16, 16}, {32, 8}};

template <class T, int S>
void space_launch(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din, const DescDev& dout,
                  const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb, const CmpSpaceDev& W,
                  int32_t* err, int part) {
  constexpr int ZL = kSpaceShapes[S][0], QB = kSpaceShapes[S][1];
  const int gb = L.cus * 4;
  if (part != kSpaceRun) {
    hipLaunchKernelGGL(k_space_clear, dim3(gb), dim3(256), 0, L.stream, W);
    hipLaunchKernelGGL(k_space_bin<false>, dim3(gb), dim3(256), 0, L.stream, sp, P, din, W, ZL, err);
    const int nb = (W.nkeys + kScanBlock - 1) / kScanBlock;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanT), 0, L.stream, W.cnt, W.nkeys, W.bsum);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kScanT), 0, L.stream, W.cnt, W.nkeys, W.bsum, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kScanT), 0, L.stream, W.cnt, W.nkeys, W.bsum);
    hipLaunchKernelGGL(k_space_bin<true>, dim3(gb), dim3(256), 0, L.stream, sp, P, din, W, ZL, err);
  }
  if (part == kSpaceBuild) return;
  if (sp.N < 0xFFFF)
    hipLaunchKernelGGL((k_cmp_space<T, ZL, QB, true>), dim3(L.cus * kColGridPerCUSpace), dim3(256), 0, L.stream, sp,
                       W, din, dout, ld, ilo, iln, newb);
  else
    hipLaunchKernelGGL((k_cmp_space<T, ZL, QB, false>), dim3(L.cus * kColGridPerCUSpace), dim3(256), 0, L.stream, sp,
                       W, din, dout, ld, ilo, iln, newb);
}

template <class T>
void space_shape(const Launch& L, int shape, const SimParams& sp, const PlanDev& P, const DescDev& din,
                 const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb,
                 const CmpSpaceDev& W, int32_t* err, int part) {
  switch (shape) {
    case 0: space_launch<T, 0>(L, sp, P, din, dout, ld, ilo, iln, newb, W, err, part); break;
    case 1: space_launch<T, 1>(L, sp, P, din, dout, ld, ilo, iln, newb, W, err, part); break;
    case 2: space_launch<T, 2>(L, sp, P, din, dout, ld, ilo, iln, newb, W, err, part); break;
    default: space_launch<T, 3>(L, sp, P, din, dout, ld, ilo, iln, newb, W, err, part);
  }
}

// workgroups per CU of a per-class comparison launch (4 wavefront tiles each);
// 512 against 64 took the grain-order comparison from 3.97 to 3.76 ms at 8000
// grains, timed on one box
constexpr int kColGridPerCU = 512;

// Shapes of the comparison tile that the autotuner chooses from per size class:
// layers per wavefront, gathers in flight per entry, and the order of a
// grain's tiles (0: column-block-major, 1: run-major).
constexpr int kCmpShapes[kCmpVariants][3] = {{8, 8, 0},  {16, 8, 0}, {16, 16, 0}, {32, 8, 0},
                                             {8, 8, 1},  {16, 8, 1}, {16, 16, 1}, {32, 8, 1}};

template <class T, int NMAX, int V>
void cmp_launch(const Launch& L, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  constexpr int ZL = kCmpShapes[V][0], QB = kCmpShapes[V][1];
  constexpr bool ZMAJ = kCmpShapes[V][2] != 0;
  hipLaunchKernelGGL((k_cmp_cols<T, NMAX, ZL, QB, ZMAJ>), dim3(tile_grid(L, kColGridPerCU, cls)), dim3(256), 0,
                     L.stream, sp, P, cls, din, dout, ld, ilo, iln, newb);
}

template <class T, int NMAX>
void cmp_variant(const Launch& L, int v, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                 const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  switch (v) {
    case 0: cmp_launch<T, NMAX, 0>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 1: cmp_launch<T, NMAX, 1>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 2: cmp_launch<T, NMAX, 2>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 3: cmp_launch<T, NMAX, 3>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 4: cmp_launch<T, NMAX, 4>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 5: cmp_launch<T, NMAX, 5>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 6: cmp_launch<T, NMAX, 6>(L, cls, sp, P, din, dout, ld, ilo, iln, newb); break;
    default: cmp_launch<T, NMAX, 7>(L, cls, sp, P, din, dout, ld, ilo, iln, newb);
  }
}

template <class T>
void comparison_class(const Launch& L, int cls, int v, const SimParams& sp, const PlanDev& P, const DescDev& din,
                      const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  switch (cls) {
    case 0: cmp_variant<T, 32>(L, v, 0, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 1: cmp_variant<T, 48>(L, v, 1, sp, P, din, dout, ld, ilo, iln, newb); break;
    case 2: cmp_variant<T, 64>(L, v, 2, sp, P, din, dout, ld, ilo, iln, newb); break;
    default: cmp_variant<T, 96>(L, v, 3, sp, P, din, dout, ld, ilo, iln, newb);
  }
}

}  // namespace

void launch_comparison_table(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                             const ListDev& ld, bool coarsen) {
  const int grid = sp.N < 65536 ? (sp.N > 0 ? sp.N : 1) : 65536;  // one wavefront per grain
  hipLaunchKernelGGL(k_cmp_table, dim3(grid), dim3(64), 0, L.stream, sp, state, din, ld, coarsen ? 1 : 0);
}

// One size class per launch (boxes n <= 32, 48, 64, 96 for cls 0..3).
void launch_comparison(const Launch& L, int cls, int variant, const SimParams& sp, const PlanDev& P,
                       const DescDev& din, const DescDev& dout, const ListDev& ld, const IdlDev& il_old,
                       const IdlDev& il_new, int32_t* newb, bool fp32) {
  if (fp32) comparison_class<float>(L, cls, variant, sp, P, din, dout, ld, il_old, il_new, newb);
  else comparison_class<double>(L, cls, variant, sp, P, din, dout, ld, il_old, il_new, newb);
}

void launch_comparison_space(const Launch& L, int shape, const SimParams& sp, const PlanDev& P, const DescDev& din,
                             const DescDev& dout, const ListDev& ld, const IdlDev& il_old, const IdlDev& il_new,
                             int32_t* newb, const CmpSpaceDev& W, int32_t* err, bool fp32, int part) {
  if (fp32) space_shape<float>(L, shape, sp, P, din, dout, ld, il_old, il_new, newb, W, err, part);
  else space_shape<double>(L, shape, sp, P, din, dout, ld, il_old, il_new, newb, W, err, part);
}

}  // namespace gg
