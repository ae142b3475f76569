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
constexpr int kGridPerCU = 512;  // workgroups per CU of a comparison launch (one per tile at 8000 grains)

constexpr int kCT = 256;
constexpr int kWaves = kCT / 64;
// LDS bytes of staged neighbour values per workgroup: FP32 keeps it small
// enough for six workgroups per CU; FP64 (VGPR-bound at five) gains more from
// fewer batches per tile.
template <class T>
constexpr int kStageBytes = sizeof(T) == 4 ? 12288 : 20480;
// Occupancy asked of the compiler: six workgroups per CU for the FP32 n <= 48
// class, which the LDS of the staged form allows (measured 6 % faster than
// the five it reaches unasked); no request elsewhere (1), where it would spill.
template <class T, int NMAX>
constexpr int kWavesPerEU = (sizeof(T) == 4 && NMAX == 48) ? 6 : 1;
constexpr int kDirectMax = 32;      // box sides up to this gather neighbour values directly
constexpr int kDepth = 2;           // list entries whose loads are in flight together (direct gathers)

typedef __attribute__((address_space(1))) const void* gptr_t;
typedef __attribute__((address_space(3))) void* lptr_t;

// Staged value at element el: one 32-bit plane for float, low and high words
// in two planes (plane stride ps) for double.
template <class T>
__device__ inline T staged(const uint32_t* sbuf, int el, int ps);
template <>
__device__ inline float staged<float>(const uint32_t* sbuf, int el, int) {
  return __uint_as_float(sbuf[el]);
}
template <>
__device__ inline double staged<double>(const uint32_t* sbuf, int el, int ps) {
  return __hiloint2double((int)sbuf[ps + el], (int)sbuf[el]);
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

// One tile = one z-layer of one grain.
//
// Items.  The comparison list is taken in passes of ITEMS entries.  An entry
// whose box covers the layer and overlaps the grain's box in x and y becomes
// an item; items are compacted in list order with the rectangle the entry
// shares with the grain's layer, the rectangle's first arena index, and per
// wavefront a mask of the rectangle's 8x8 voxel blocks that wavefront owns.
//
// Boxes up to kDirectMax skip the items and the staging: their wavefronts
// gather the neighbour values straight from the arena (described at that
// branch), which keeps the workgroup free of LDS and of barriers until the
// bounding-box reduction.
//
// Staging.  Each item's rectangle is staged row-major in LDS, padded to whole
// 64-element segments, the items following each other (a_c0 = first segment
// of each).  The segments are copied in batches of NCH by direct global-to-LDS
// loads: one wavefront load fills one segment (one 32-bit plane of it), so the
// item, its arena index and its width are uniform across the wavefront, and
// the loads hold no registers while in flight.
//
// Maximum.  The layer is cut into 8x8 blocks, block b owned by wavefront b % 4,
// one voxel per lane and block, so a lane's voxels stay in its registers for
// the whole tile.  For each item of the batch, in list order, a wavefront
// visits only the blocks the rectangle touches (one scalar bit test per block)
// and updates with a strict '>'.  A rectangle may
// straddle two batches: a voxel's staged index grows with the entry's list
// position, so each voxel still sees its neighbours in list order.
template <class T, int NMAX>
__global__ __launch_bounds__(kCT) __attribute__((amdgpu_waves_per_eu(kWavesPerEU<T, NMAX>))) void k_comparison(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                    ListDev ld, IdlDev ilo, IdlDev iln, int32_t* newb) {
  constexpr int NB = (NMAX + 7) / 8;                    // blocks per side, at most
  constexpr int VPT = (NB * NB + kWaves - 1) / kWaves;  // blocks per wavefront = voxels per lane
  constexpr bool kDirect = NMAX <= kDirectMax;          // gathers straight from the arena, no staging
  static_assert(VPT <= 64, "a wavefront's block mask is one 64-bit word");
  constexpr int ITEMS = 128;                            // list entries per pass (at most 256: a_of is 8-bit)
  static_assert(ITEMS % 64 == 0 && ITEMS <= kCT, "whole wavefronts compact the items");
  constexpr int NCH = kStageBytes<T> / (int)sizeof(T) / 64;  // staged segments per batch
  constexpr int SB = NCH / kWaves;                          // segments per wavefront and batch
  constexpr int PL = (int)sizeof(T) / 4;                    // 32-bit planes of a staged value
  static_assert(SB * kWaves == NCH, "a batch is whole segments for every wavefront");
  __shared__ uint32_t sbuf[PL * NCH * 64];  // plane p of segment c, lane l at [p*NCH*64 + c*64 + l]
  __shared__ int4 a_stg[ITEMS];  // arena index (lo, hi), ceil(2^32 / width), nn | width << 8 | height << 16
  __shared__ uint64_t a_wm[kWaves][ITEMS];  // blocks of the rectangle owned by each wavefront
  __shared__ int32_t a_id[ITEMS], a_c0[ITEMS + 1];
  __shared__ uint32_t a_rect[ITEMS];  // local rectangle x0 | x1 << 8 | y0 << 16 | y1 << 24
  __shared__ uint8_t a_of[NCH];       // item of each segment of the batch
  __shared__ int s_cnt[ITEMS / 64], s_seg[ITEMS / 64];
  __shared__ int s_red[6][kWaves];
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const T lim = (T)(0.7 * sp.delta);
  const int maxL = ld.max_list;
  const TileRange tr = xcd_tiles(total);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], x1 = b[3], y1 = b[4];
    const int n = x1 - x0, nn2 = n * n;
    const int nbx = (n + 7) >> 3;
    const int k = z0 + kl;
    const int4* tab = ld.tab + (int64_t)g * maxL;  // neighbours overlapping the box, in list order
    const int32_t* tid_of = ld.tab_id + (int64_t)g * maxL;
    const int len = ld.tab_len[g];

    // The lane's voxel in its r-th block: (i - y0, j - x0) = (by*8 + ly, bx*8 + lx)
    // with the block origin uniform across the wavefront.
    const int ly = lane >> 3, lx = lane & 7;
    const uint32_t nbx_inv = (65536u + (uint32_t)nbx - 1u) / (uint32_t)nbx;  // bb / nbx = bb * nbx_inv >> 16 (bb < 2^12)
    auto block_origin = [&](int r, int& oy, int& ox) {
      const int bb = wv + kWaves * r;
      const int by = (int)(((uint32_t)bb * nbx_inv) >> 16);
      oy = by * 8;
      ox = (bb - by * nbx) * 8;
      return bb < nbx * nbx;
    };
    T best[VPT];
    uint32_t bid[VPT];
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      best[r] = (T)(-1.0);
      bid[r] = 0xFFFFFFFFu;  // not covered
    }

    if constexpr (!kDirect) {
      for (int lbase = 0; lbase < len; lbase += ITEMS) {
        // compact the items of this pass, keeping list order
        const int idx = lbase + tid;
        bool keep = false;
        int nseg = 0, nb = 0, nn = 0, rx0 = 0, rx1 = 0, ry0 = 0, ry1 = 0;
        int64_t gbase = 0;
        if (tid < ITEMS && idx < len) {
          const int4 e = tab[idx];
          nb = tid_of[idx];
          const int q2 = e.w & 0xFFFF, q5 = (int)((uint32_t)e.w >> 16);
          if (q2 <= k && k < q5) {
            keep = true;
            nn = (int)((uint32_t)e.y >> 16);
            const uint32_t rc = (uint32_t)e.z;
            rx0 = rc & 0xFF; rx1 = (rc >> 8) & 0xFF; ry0 = (rc >> 16) & 0xFF; ry1 = rc >> 24;
            nseg = ((rx1 - rx0) * (ry1 - ry0) + 63) >> 6;
            gbase = (int64_t)((uint64_t)((uint32_t)e.y & 0xFFFFu) << 32 | (uint32_t)e.x) + (int64_t)(k - q2) * nn * nn;
          }
        }
        const uint64_t m = __ballot(keep);
        int sc = nseg;  // inclusive scan of the segment counts over the wavefront
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(sc, off);
          if (lane >= off) sc += y;
        }
        if (lane == 63 && wv < ITEMS / 64) {
          s_cnt[wv] = __popcll(m);
          s_seg[wv] = sc;
        }
        __syncthreads();
        if (keep) {
          int pos = __popcll(m & ((1ull << lane) - 1ull)), c0 = sc - nseg;
          for (int w = 0; w < wv; ++w) {
            pos += s_cnt[w];
            c0 += s_seg[w];
          }
          uint64_t wm[kWaves] = {};
          for (int by = ry0 >> 3; by <= (ry1 - 1) >> 3; ++by)
            for (int bx = rx0 >> 3; bx <= (rx1 - 1) >> 3; ++bx) {
              const int bb = by * nbx + bx;
              wm[bb % kWaves] |= 1ull << (bb / kWaves);
            }
          for (int w = 0; w < kWaves; ++w) a_wm[w][pos] = wm[w];
          const int rw = rx1 - rx0, rh = ry1 - ry0;
          a_id[pos] = nb;
          a_rect[pos] = (uint32_t)rx0 | (uint32_t)rx1 << 8 | (uint32_t)ry0 << 16 | (uint32_t)ry1 << 24;
          a_stg[pos] = make_int4((int)(uint32_t)(uint64_t)gbase, (int)(uint32_t)((uint64_t)gbase >> 32),
                                 rw > 1 ? (int)(0xFFFFFFFFu / (uint32_t)rw + 1u) : 0, nn | rw << 8 | rh << 16);
          a_c0[pos] = c0;
        }
        int cnt = 0, C = 0;  // items and staged segments of this pass
        for (int w = 0; w < ITEMS / 64; ++w) {
          cnt += s_cnt[w];
          C += s_seg[w];
        }
        if (tid == 0) a_c0[cnt] = C;
        __syncthreads();

        for (int cs = 0; cs < C; cs += NCH) {
          const int nc = min(NCH, C - cs);
          if (tid < nc) {  // item of each segment: last a with a_c0[a] <= cs + tid
            int a = 0;
  #pragma unroll
            for (int st = ITEMS / 2; st > 0; st >>= 1)
              if (a + st < cnt && a_c0[a + st] <= cs + tid) a += st;
            a_of[tid] = (uint8_t)a;
          }
          __syncthreads();
          // one direct global-to-LDS load per segment and plane: lane l of the
          // wavefront writes word l of the segment, lanes past the rectangle's end
          // load its first element into the padding
  #pragma unroll
          for (int u = 0; u < SB; ++u) {
            const int c = u * kWaves + wv;  // this wavefront's segment (uniform)
            if (c < nc) {
              const int a = __builtin_amdgcn_readfirstlane((int)a_of[c]);
              const int4 e4 = a_stg[a];
              const uint32_t blo = (uint32_t)__builtin_amdgcn_readfirstlane(e4.x);
              const uint32_t bhi = (uint32_t)__builtin_amdgcn_readfirstlane(e4.y);
              const uint32_t mg = (uint32_t)__builtin_amdgcn_readfirstlane(e4.z);
              const int wd = __builtin_amdgcn_readfirstlane(e4.w);
              const int c0 = __builtin_amdgcn_readfirstlane(a_c0[a]);
              const int nn_ = wd & 0xFF, rw = (wd >> 8) & 0xFF, rh = wd >> 16;
              const int r = ((cs + c - c0) << 6) + lane;  // element of the rectangle
              const int qr = rw > 1 ? (int)__umulhi((uint32_t)r, mg) : r;  // r / rw
              const int64_t off = qr < rh ? (int64_t)qr * nn_ + (r - qr * rw) : 0;
              const char* src = (const char*)(ain + ((int64_t)((uint64_t)bhi << 32 | blo) + off));
  #pragma unroll
              for (int pl = 0; pl < PL; ++pl)
                __builtin_amdgcn_global_load_lds((gptr_t)(src + 4 * pl), (lptr_t)(sbuf + pl * NCH * 64 + c * 64), 4,
                                                 0, 0);
            }
          }
          __builtin_amdgcn_s_waitcnt(0);  // this wavefront's LDS writes are done before the barrier
          __syncthreads();
          const uint32_t lim_e = (uint32_t)(nc << 6);
          // the items with segments in the batch, from the item of its first segment
          for (int aa = a_of[0]; aa < cnt && a_c0[aa] < cs + nc; ++aa) {
            const uint64_t wmask = ((uint64_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a_wm[wv][aa] >> 32)) << 32) |
                                   (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a_wm[wv][aa]);
            const uint32_t rc = a_rect[aa];
            const int rx0 = rc & 0xFF, ry0 = (rc >> 16) & 0xFF;
            const int rw = (int)((rc >> 8) & 0xFFu) - rx0, rh = (int)(rc >> 24) - ry0;
            const int pre = (a_c0[aa] - cs) << 6;  // sbuf index of the rectangle's first element (may be < 0)
            const uint32_t id = (uint32_t)a_id[aa];
  #pragma unroll
            for (int r = 0; r < VPT; ++r) {
              if ((wmask >> r) & 1ull) {
                int oy, ox;
                block_origin(r, oy, ox);
                const int dx = lx + (ox - rx0), dy = ly + (oy - ry0);
                const int el = pre + dy * rw + dx;
                const bool ok = (uint32_t)dx < (uint32_t)rw && (uint32_t)dy < (uint32_t)rh && (uint32_t)el < lim_e;
                const T val = staged<T>(sbuf, ok ? el : 0, NCH * 64);
                const bool up = ok && val > best[r];
                best[r] = up ? val : best[r];
                bid[r] = up ? id : bid[r];
              }
            }
          }
          __syncthreads();
        }
      }
    } else {
      // Direct gathers (boxes up to kDirectMax): each wavefront walks the list on
      // its own, one entry per lane to set up a pass of 64 entries (the shared
      // rectangle, and where the entry's layer k starts), compacts the entries
      // that touch the layer with a ballot and takes them in list order,
      // kDepth at a time: the loads of all their blocks are issued before the
      // first update.  Lanes off a rectangle load its first voxel and do not update.
      int yr[VPT], xr[VPT];  // the lane's voxel in block r, local to the box
#pragma unroll
      for (int r = 0; r < VPT; ++r) {
        int oy, ox;
        block_origin(r, oy, ox);
        yr[r] = oy + ly;
        xr[r] = ox + lx;
      }
      int boy = 0, box = 0;  // lane r < VPT: origin of block r, to test all blocks at once
      bool bin = false;
#pragma unroll
      for (int r = 0; r < VPT; ++r)
        if (lane == r) bin = block_origin(r, boy, box);
      for (int pb = 0; pb < len; pb += 64) {
        const int idx = pb + lane;
        bool keep = false;
        int nb = 0, nn = 0, shift = 0, first = 0;
        uint32_t rect = 0;
        int64_t layer = 0;
        if (idx < len) {
          const int4 e = tab[idx];
          nb = tid_of[idx];
          const int q2 = e.w & 0xFFFF, q5 = (int)((uint32_t)e.w >> 16);
          if (q2 <= k && k < q5) {
            keep = true;
            nn = (int)((uint32_t)e.y >> 16);
            rect = (uint32_t)e.z;
            // the rectangle's first voxel in the neighbour's layer k; voxel (il, jl)
            // of the grain's layer is at il*nn + jl + shift from there
            layer = (int64_t)((uint64_t)((uint32_t)e.y & 0xFFFFu) << 32 | (uint32_t)e.x) + (int64_t)(k - q2) * nn * nn;
            shift = -(int)((rect >> 16) & 0xFF) * nn - (int)(rect & 0xFF);
            first = 0;
          }
        }
        uint64_t m = __ballot(keep);
        while (m) {
          T val[kDepth][VPT];
          bool ok[kDepth][VPT];
          uint64_t ov[kDepth];
          uint32_t id[kDepth];
#pragma unroll
          for (int d = 0; d < kDepth; ++d) {
            ov[d] = 0;
            id[d] = 0;
            if (m == 0) continue;
            const int a = (int)__builtin_ctzll(m);  // next entry in list order
            m &= m - 1;
            id[d] = (uint32_t)__builtin_amdgcn_readlane(nb, a);
            const int nn_a = __builtin_amdgcn_readlane(nn, a);
            const int shift_a = __builtin_amdgcn_readlane(shift, a);
            const uint32_t first_a = (uint32_t)__builtin_amdgcn_readlane(first, a);
            const uint32_t rc = (uint32_t)__builtin_amdgcn_readlane((int)rect, a);
            const uint32_t llo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)layer, a);
            const uint32_t lhi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)layer >> 32), a);
            const T* __restrict__ src = ain + (int64_t)((uint64_t)lhi << 32 | llo);
            const int rx0 = rc & 0xFF, rx1 = (rc >> 8) & 0xFF, ry0 = (rc >> 16) & 0xFF, ry1 = rc >> 24;
            const int rw = rx1 - rx0, rh = ry1 - ry0;
            // blocks the rectangle touches, and those it covers whole
            ov[d] = __ballot(bin & (box < rx1) & (box + 8 > rx0) & (boy < ry1) & (boy + 8 > ry0));
            const uint64_t fm = __ballot(bin & (box >= rx0) & (box + 8 <= rx1) & (boy >= ry0) & (boy + 8 <= ry1));
#pragma unroll
            for (int r = 0; r < VPT; ++r) {
              ok[d][r] = true;
              val[d][r] = (T)0;
              if ((ov[d] >> r) & 1ull) {
                uint32_t off = (uint32_t)(__mul24(yr[r], nn_a) + xr[r] + shift_a);  // 24-bit product: both < 256
                if (!((fm >> r) & 1ull)) {
                  ok[d][r] = ((uint32_t)(xr[r] - rx0) < (uint32_t)rw) & ((uint32_t)(yr[r] - ry0) < (uint32_t)rh);
                  off = ok[d][r] ? off : first_a;
                }
                val[d][r] = src[off];
              }
            }
          }
#pragma unroll
          for (int d = 0; d < kDepth; ++d) {
#pragma unroll
            for (int r = 0; r < VPT; ++r) {
              if ((ov[d] >> r) & 1ull) {
                const bool up = ok[d][r] & (val[d][r] > best[r]);
                best[r] = up ? val[d][r] : best[r];
                bid[r] = up ? id[d] : bid[r];
              }
            }
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
    const int64_t lbase = (int64_t)kl * nn2;
    const T* own = ain + din.off[g] + lbase;
    T* dst = aout + dout.off[g] + lbase;
    // maxima start at -INT_MAX so that their negation (reduced with min below) cannot overflow
    int bx0 = INT_MAX, by0 = INT_MAX, bz0 = INT_MAX, bx1 = -INT_MAX, by1 = -INT_MAX, bz1 = -INT_MAX;
#pragma unroll
    for (int r = 0; r < VPT; ++r) {
      int oy, ox;
      const bool inb = block_origin(r, oy, ox);
      const int il = oy + ly, jl = ox + lx;
      if (!inb || il >= n || jl >= n) continue;
      const int i = y0 + il, j = x0 + jl;
      const int v = il * n + jl;
      if (wall && !deep_inside(j, i, k, sp.gb, sp.ngp)) {
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

template <class T>
void comparison_class(const Launch& L, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                      const DescDev& dout, const ListDev& ld, const IdlDev& ilo, const IdlDev& iln, int32_t* newb) {
  switch (cls) {
    case 0:
      hipLaunchKernelGGL((k_comparison<T, 32>), dim3(tile_grid(L, kGridPerCU, 0)), dim3(kCT), 0, L.stream, sp, P, 0, din, dout, ld, ilo,
                         iln, newb);
      break;
    case 1:
      hipLaunchKernelGGL((k_comparison<T, 48>), dim3(tile_grid(L, kGridPerCU, 1)), dim3(kCT), 0, L.stream, sp, P, 1, din, dout, ld, ilo,
                         iln, newb);
      break;
    case 2:
      hipLaunchKernelGGL((k_comparison<T, 64>), dim3(tile_grid(L, kGridPerCU, 2)), dim3(kCT), 0, L.stream, sp, P, 2, din, dout, ld, ilo,
                         iln, newb);
      break;
    default:
      hipLaunchKernelGGL((k_comparison<T, 96>), dim3(tile_grid(L, kGridPerCU, 3)), dim3(kCT), 0, L.stream, sp, P, 3, din, dout, ld, ilo,
                         iln, newb);
  }
}

}  // namespace

void launch_comparison_table(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                             const ListDev& ld, bool coarsen) {
  const int grid = sp.N < 65536 ? (sp.N > 0 ? sp.N : 1) : 65536;  // one wavefront per grain
  hipLaunchKernelGGL(k_cmp_table, dim3(grid), dim3(64), 0, L.stream, sp, state, din, ld, coarsen ? 1 : 0);
}

// One size class per launch (boxes n <= 32, 48, 64, 96 for cls 0..3).
void launch_comparison(const Launch& L, int cls, const SimParams& sp, const PlanDev& P, const DescDev& din,
                       const DescDev& dout, const ListDev& ld, const IdlDev& il_old, const IdlDev& il_new,
                       int32_t* newb, bool fp32) {
  if (fp32) comparison_class<float>(L, cls, sp, P, din, dout, ld, il_old, il_new, newb);
  else comparison_class<double>(L, cls, sp, P, din, dout, ld, il_old, il_new, newb);
}

}  // namespace gg
