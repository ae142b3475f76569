// Level-set stage: grain existence, volume and surface area from the marching-
// cubes hull of the compared distance field, deletion of vanished grains and
// the box resize (grainhdl::level_set, src/grainhdl.cpp:669-684 ->
// LSbox::extractContour, src/box.cpp:1144-1161).
//
// k_hull walks the cells of one z-layer of a grain's input box exactly as
// MarchingCubesAlgorithm::generateHull (src/marchingCubes.cpp:364-478): case
// index from the eight corners (inside = phi >= 0), VertexInterp
// (:484-499) on the cut edges, triangles from the case table, then the
// signed-tetrahedron volume term (src/grainHull.cpp:40-56, including its
// float(1/6) factor) and the triangle area (:120-131).  A grain exists iff
// some cell is cut.  Each tile keeps Neumaier-compensated FP64 sums and writes
// them to its own slot; k_finalize adds a grain's slots in layer order, so the
// result does not depend on scheduling.
#include <algorithm>

#include "gg_internal.h"

namespace gg {
namespace {
constexpr int kGridPerCU = 64;  // workgroups per CU of a hull launch (each loads the triangle table once)
#include "mc_cases.inc"

// Triangle table row of each case: edge indices in nibbles, first triangle in
// the low nibbles, 0xF after the last entry.
__constant__ uint64_t c_tri[256];

// Cube corner q -> (dx,dy,dz) (src/marchingCubes.cpp:378-386) as bit masks
// over q: {0,0,1} {1,0,1} {1,0,0} {0,0,0} {0,1,1} {1,1,1} {1,1,0} {0,1,0}.
constexpr unsigned kCornerX = 0x66u, kCornerY = 0xF0u, kCornerZ = 0x33u;
// Edge e -> corner pair (src/marchingCubes.cpp:393-464) in nibbles:
// {0,1} {1,2} {2,3} {3,0} {4,5} {5,6} {6,7} {7,4} {0,4} {1,5} {2,6} {3,7}.
constexpr uint64_t kEdgeA = 0x321076543210ull, kEdgeB = 0x765447650321ull;

struct Acc {
  double s, c;
  __device__ void add(double x) {
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ void merge(double s2, double c2) {
    double t = s + s2;
    double bp = t - s;
    double e = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 + e;
  }
};

constexpr int kHT = 256;
constexpr int kTopoLdsBits = 7, kTopoLds = 1 << kTopoLdsBits;  // per-tile topology table
constexpr int kTopoCapBits = 6;
static_assert((1 << kTopoCapBits) == kTopoCap, "grain topology table size");
static_assert(kTopoCap <= 64, "a junction list is a 64-bit mask over the grain table's slots");
constexpr int kJuncLdsBits = 8, kJuncLds = 1 << kJuncLdsBits;  // per-tile junction table
constexpr int kJuncCapBits = 9;
static_assert((1 << kJuncCapBits) == kJuncCap, "grain junction table size");
constexpr int kCPT = 4;              // cells classified per lane and pass (a 4 KiB queue: six workgroups per CU)
constexpr int kChunk = kCPT * kHT;  // cells classified per pass

// Offset of corner q of the cell (xl, yl) in the two staged layers L[0..2*NMAX^2).
template <int NMAX>
__device__ inline int corner_off(int q, int xl, int yl, int n) {
  return (int)((kCornerZ >> q) & 1u) * (NMAX * NMAX) + (yl + (int)((kCornerY >> q) & 1u)) * n + xl +
         (int)((kCornerX >> q) & 1u);
}

// VertexInterp (src/marchingCubes.cpp:484-499) at iso level 0 on cube edge e of
// the cell with lower corner (x,y,z); the corner values come from LDS.
template <class T, int NMAX>
__device__ inline void edge_vertex(int e, const T* L, int xl, int yl, int n, int x, int y, int z, double& px,
                                   double& py, double& pz) {
  const int a = (int)((kEdgeA >> (4 * e)) & 15u), b = (int)((kEdgeB >> (4 * e)) & 15u);
  const double v1 = (double)L[corner_off<NMAX>(a, xl, yl, n)];
  const double v2 = (double)L[corner_off<NMAX>(b, xl, yl, n)];
  const double p1x = x + (int)((kCornerX >> a) & 1u), p1y = y + (int)((kCornerY >> a) & 1u),
               p1z = z + (int)((kCornerZ >> a) & 1u);
  const double p2x = x + (int)((kCornerX >> b) & 1u), p2y = y + (int)((kCornerY >> b) & 1u),
               p2z = z + (int)((kCornerZ >> b) & 1u);
  if (fabs(v1) < 0.00001) {
    px = p1x; py = p1y; pz = p1z;
  } else if (fabs(v2) < 0.00001) {
    px = p2x; py = p2y; pz = p2z;
  } else if (fabs(v1 - v2) < 0.00001) {
    px = p1x; py = p1y; pz = p1z;
  } else {
    // An edge runs along one axis, so p2 - p1 is a unit step and only mu is
    // rounded. The FP32 path takes mu in FP32 (reciprocal, ~2 ulp): the values
    // it interpolates are FP32 already; the FP64 path divides in FP64.
    double mu;
    if constexpr (sizeof(T) == 4) {
      const float f1 = (float)v1, f2 = (float)v2;
      mu = (double)(f1 * __builtin_amdgcn_rcpf(f1 - f2));
    } else {
      mu = v1 / (v1 - v2);
    }
    px = p1x + mu * (p2x - p1x);
    py = p1y + mu * (p2y - p1y);
    pz = p1z + mu * (p2z - p1z);
  }
}

// One (grain, z-layer) tile: layers kl and kl+1 are staged in LDS; each lane
// classifies kCPT cells, the triangles of the cut ones (some corners inside,
// some outside) are listed in a fixed order by one workgroup scan, and then
// every lane works on one triangle at a time, so the triangle work is neither
// left to a few lanes of each wavefront nor held up by cells with many
// triangles.  The triangle table sits in LDS; corner and edge tables are bit
// masks, so no lookup goes to memory with a lane-dependent address.
// FP32: six waves per SIMD asked of the compiler, which the LDS of both
// classes then allows.
//
// TOPO: each cut cell also names its grains (the grain itself at inside
// corners, the IDLocal id at the others), a per-tile LDS table keeps the first
// appearance of each id, and the tile adds it to the grain's table.  A
// separate instantiation, so the registers it needs cost nothing when off.
// A cell's barycenter sum (non-negative grid coordinates) in fixed point:
// integer adds make a list's position sum independent of the order in which
// tiles and lanes add to it (exact to 2^-33 grid units per cell).
__device__ inline unsigned long long fixpos(double v) {
  return (unsigned long long)__double2ll_rn(v * kPosScale);
}

template <class T, int NMAX, bool TOPO>
__global__ __launch_bounds__(kHT) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1)))
void k_hull(SimParams sp, PlanDev P, int cls, DescDev din, int32_t* exflag, double* partials, IdlDev il,
            TopoDev topo, int32_t* err) {
  __shared__ T L[2 * NMAX * NMAX];
  __shared__ uint32_t q[kChunk];
  __shared__ uint64_t s_tri[256];
  __shared__ uint8_t s_ntri[256];  // triangles of each case
  __shared__ int s_wave[kHT / 64];
  __shared__ double s_red[4][kHT / 64];
  __shared__ int s_any[kHT / 64];
  __shared__ int s_tk[kTopoLds], s_to[kTopoLds];  // per-tile topology table (id, first appearance)
  __shared__ unsigned long long s_jm[kJuncLds];    // per-tile junction lists (slot masks)
  __shared__ int s_jg[TOPO ? kJuncLds : 1];        // the grain-table slot each per-tile list went to
  // Element sums per tile list in LDS, added to the grain's table once per tile
  // and list; the FP64 n <= 96 class has no LDS left for them and adds each
  // cell's sums to the grain's table directly.
  constexpr bool kElemLds = TOPO && !(sizeof(T) == 8 && NMAX == 96);
  __shared__ unsigned long long s_jp[kElemLds ? 3 * kJuncLds : 1];  // fixed point, kPosScale per grid unit
  __shared__ double s_ja[kElemLds ? kJuncLds : 1];  // summed triangle areas
  __shared__ int s_jc[kElemLds ? kJuncLds : 1];
  __shared__ int s_gk[kTopoCap];                   // copy of the grain's id table
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* arena = (const T*)din.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < 256; i += kHT) {
    uint64_t row = c_tri[i];
    s_tri[i] = row;
    int cnt = 0;
    for (; (row & 15u) != 15u; row >>= 12) ++cnt;
    s_ntri[i] = (uint8_t)cnt;
  }
  if constexpr (TOPO) {
    for (int i = tid; i < kTopoLds; i += kHT) {
      s_tk[i] = -1;
      s_to[i] = INT_MAX;
    }
    for (int i = tid; i < kJuncLds; i += kHT) {
      s_jm[i] = 0ull;
      if constexpr (kElemLds) {
        s_jc[i] = 0;
        s_ja[i] = 0.0;
        s_jp[3 * i] = s_jp[3 * i + 1] = s_jp[3 * i + 2] = 0ull;
      }
    }
  }
  // consecutive tiles (the layers of one grain) on one XCD: the layer two
  // tiles share, and the grain's tables, come from that XCD's L2
  const TileRange tr = xcd_tiles(total);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int64_t slot = P.layer_base[cls] + tile;
    const int32_t* b = din.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2];
    const int n = b[3] - b[0], nn2 = n * n, m1 = n - 1;
    const int z = z0 + kl;
    // c / m1 = c * m1_inv >> 24 for every cell index c < m1^2 (m1 < 96)
    const uint32_t m1_inv = m1 > 0 ? ((1u << 24) + (uint32_t)m1 - 1u) / (uint32_t)m1 : 0u;
    Acc vol{0.0, 0.0}, area{0.0, 0.0};
    int any = 0;
    if (kl < m1) {
      const T* box = arena + din.off[g] + (int64_t)kl * nn2;
      stage_to_lds(L, box, nn2, tid, kHT);
      stage_to_lds(L + NMAX * NMAX, box + nn2, nn2, tid, kHT);
      __syncthreads();
      for (int c0 = 0; c0 < m1 * m1; c0 += kChunk) {
        uint32_t code[kCPT];
        int mine = 0;  // triangles of this lane's cut cells
#pragma unroll
        for (int r = 0; r < kCPT; ++r) {
          const int c = c0 + r * kHT + tid;
          code[r] = 0u;
          if (c < m1 * m1) {
            const int yl = (int)(((uint32_t)c * m1_inv) >> 24), xl = c - yl * m1;
            unsigned ci = 0;
#pragma unroll
            for (int qq = 0; qq < 8; ++qq)
              if (L[corner_off<NMAX>(qq, xl, yl, n)] >= T(0)) ci |= 1u << qq;
            if (ci != 0u && ci != 255u) {
              code[r] = ci | ((uint32_t)yl << 8) | ((uint32_t)xl << 16);
              mine += s_ntri[ci];
            }
          }
        }
        if constexpr (TOPO) {
          const int32_t* ib = il.bounds + 6 * g;
          const int ix0 = ib[0], iy0 = ib[1], iz0 = ib[2], inx = ib[3] - ib[0], iny = ib[4] - ib[1],
                    inz = ib[5] - ib[2];
          const uint32_t* idl = il.arena + il.off[g];
          // the ids a cut cell names (own id inside, IDLocal id outside, none
          // outside the IDLocal box) and which of them are first occurrences
          auto cell_ids = [&](uint32_t cd, uint32_t (&ids)[8], bool (&first)[8]) {
            const unsigned ci = cd & 0xFFu;
            const int yl = (cd >> 8) & 0xFF, xl = cd >> 16;
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
              uint32_t id = (uint32_t)g;
              if (!((ci >> qq) & 1u)) {
                const int cx = x0 + xl + (int)((kCornerX >> qq) & 1u) - ix0;
                const int cy = y0 + yl + (int)((kCornerY >> qq) & 1u) - iy0;
                const int cz = z + (int)((kCornerZ >> qq) & 1u) - iz0;
                id = ((unsigned)cx < (unsigned)inx && (unsigned)cy < (unsigned)iny && (unsigned)cz < (unsigned)inz)
                         ? idl[((int64_t)cz * iny + cy) * inx + cx]
                         : 0xFFFFFFFFu;
              }
              ids[qq] = id;
            }
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
              first[qq] = true;
#pragma unroll
              for (int q2 = 0; q2 < qq; ++q2) first[qq] &= ids[q2] != ids[qq];
            }
          };
          // the barycenter sums (src/triangle.h:12-18) and number of a cut
          // cell's triangles, vertices as the hull's
          auto cell_elems = [&](uint32_t cd, double& sx, double& sy, double& sz, double& sa) -> int {
            const unsigned ci = cd & 0xFFu;
            const int cyl = (cd >> 8) & 0xFF, cxl = (cd >> 16) & 0xFF;
            const int nt = s_ntri[ci];
            uint64_t row = s_tri[ci];
            sx = sy = sz = sa = 0.0;
            for (int j = 0; j < nt; ++j, row >>= 12) {
              double p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
              edge_vertex<T, NMAX>((int)(row & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p0x, p0y, p0z);
              edge_vertex<T, NMAX>((int)((row >> 4) & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p1x, p1y, p1z);
              edge_vertex<T, NMAX>((int)((row >> 8) & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p2x, p2y, p2z);
              sx += (p0x + p1x + p2x) / 3;
              sy += (p0y + p1y + p2y) / 3;
              sz += (p0z + p1z + p2z) / 3;
              // GrainBoundary::computeArea (src/InterfacialElement.cpp:81-92)
              const double abx = p0x - p1x, aby = p0y - p1y, abz = p0z - p1z;
              const double acx = p0x - p2x, acy = p0y - p2y, acz = p0z - p2z;
              const double cx = aby * acz - abz * acy, cy = abz * acx - abx * acz, cz = abx * acy - aby * acx;
              sa += sqrt(cx * cx + cy * cy + cz * cz) / 2.0;
            }
            return nt;
          };
          // 1. grains: each distinct id with its first-appearance key (cell in the
          //    (z, y, x) loop * 8 + rank among the cell's ascending ids), per tile
          //    in LDS, then into the grain's table
#pragma unroll
          for (int r = 0; r < kCPT; ++r) {
            if (code[r] == 0u) continue;
            uint32_t ids[8];
            bool first[8];
            cell_ids(code[r], ids, first);
            const int cell = (kl * m1 + (int)((code[r] >> 8) & 0xFF)) * m1 + (int)(code[r] >> 16);
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
              if (!first[qq] || ids[qq] == 0xFFFFFFFFu) continue;
              int rank = 0;
#pragma unroll
              for (int q2 = 0; q2 < 8; ++q2) rank += (first[q2] & (ids[q2] < ids[qq])) ? 1 : 0;
              const int key = (int)ids[qq], o = cell * 8 + rank;
              uint32_t hsl = ((uint32_t)key * 2654435761u) >> (32 - kTopoLdsBits);
              bool done = false;
              for (int pr = 0; pr < kTopoLds && !done; ++pr) {
                const int old = atomicCAS(&s_tk[hsl], -1, key);
                if (old == -1 || old == key) {
                  atomicMin(&s_to[hsl], o);
                  done = true;
                }
                hsl = (hsl + 1) & (kTopoLds - 1);
              }
              if (!done) atomicOr(err, kErrTopo);
            }
          }
          __syncthreads();
          int32_t* gk = topo.key + (int64_t)g * kTopoCap;
          int32_t* go = topo.ord + (int64_t)g * kTopoCap;
          for (int i = tid; i < kTopoLds; i += kHT) {
            const int key = s_tk[i];
            if (key != -1) {
              uint32_t hsl = ((uint32_t)key * 2654435761u) >> (32 - kTopoCapBits);
              bool done = false;
              for (int pr = 0; pr < kTopoCap && !done; ++pr) {
                const int old = atomicCAS(&gk[hsl], -1, key);
                if (old == -1 || old == key) {
                  atomicMin(&go[hsl], s_to[i]);
                  done = true;
                }
                hsl = (hsl + 1) & (kTopoCap - 1);
              }
              if (!done) atomicOr(err, kErrTopo);
            }
            s_tk[i] = -1;
            s_to[i] = INT_MAX;
          }
          __syncthreads();
          // the grain's id table as it stands now (this chunk's ids are in it), read
          // past the L1 once, for the slot lookups below
          for (int i = tid; i < kTopoCap; i += kHT)
            s_gk[i] = __hip_atomic_load(&gk[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __syncthreads();
          // 2. junctions: a cell's list of distinct ids as a mask over the slots
          //    its ids hold in the grain's table (a slot never changes once
          //    taken, so equal lists give equal masks); distinct masks per tile in
          //    LDS, then into the grain's junction table
          uint64_t cmask[kCPT];
          int cjsl[kCPT];  // the per-tile list of each cut cell
#pragma unroll
          for (int r = 0; r < kCPT; ++r) {
            cmask[r] = 0;
            cjsl[r] = -1;
            if (code[r] == 0u) continue;
            uint32_t ids[8];
            bool first[8];
            cell_ids(code[r], ids, first);
            uint64_t mask = 0;
#pragma unroll
            for (int qq = 0; qq < 8; ++qq) {
              if (!first[qq] || ids[qq] == 0xFFFFFFFFu) continue;
              uint32_t hsl = (ids[qq] * 2654435761u) >> (32 - kTopoCapBits);
              for (int pr = 0; pr < kTopoCap; ++pr) {
                const int k = s_gk[hsl];
                if (k == (int)ids[qq]) {
                  mask |= 1ull << hsl;
                  break;
                }
                if (k == -1) break;  // not in the table (it was full: reported above)
                hsl = (hsl + 1) & (kTopoCap - 1);
              }
            }
            cmask[r] = mask;
            if (mask == 0) continue;
            uint32_t hsl = (uint32_t)((mask * 0x9E3779B97F4A7C15ull) >> (64 - kJuncLdsBits));
            int jsl = -1;
            for (int pr = 0; pr < kJuncLds && jsl < 0; ++pr) {
              const unsigned long long old = atomicCAS(&s_jm[hsl], 0ull, (unsigned long long)mask);
              if (old == 0ull || old == mask) jsl = (int)hsl;
              hsl = (hsl + 1) & (kJuncLds - 1);
            }
            if (jsl < 0) {
              atomicOr(err, kErrTopo);
              continue;
            }
            cjsl[r] = jsl;
            if constexpr (kElemLds) {
              // 3. interfacial elements (src/grainHull.cpp:227-270), per tile list
              double sx, sy, sz, sa;
              const int nt = cell_elems(code[r], sx, sy, sz, sa);
              atomicAdd(&s_jp[3 * jsl], fixpos(sx));
              atomicAdd(&s_jp[3 * jsl + 1], fixpos(sy));
              atomicAdd(&s_jp[3 * jsl + 2], fixpos(sz));
              atomicAdd(&s_ja[jsl], sa);
              atomicAdd(&s_jc[jsl], nt);
            }
          }
          __syncthreads();
          unsigned long long* gj = topo.jkey + (int64_t)g * kJuncCap;
          unsigned long long* gp = topo.jpos + (int64_t)g * kJuncCap * 3;
          int32_t* gc = topo.jcnt + (int64_t)g * kJuncCap;
          double* ga = topo.jarea + (int64_t)g * kJuncCap;
          for (int i = tid; i < kJuncLds; i += kHT) {
            const unsigned long long m = s_jm[i];
            if (m != 0ull) {
              uint32_t hsl = (uint32_t)((m * 0x9E3779B97F4A7C15ull) >> (64 - kJuncCapBits));
              int gsl = -1;
              for (int pr = 0; pr < kJuncCap && gsl < 0; ++pr) {
                const unsigned long long old = atomicCAS(&gj[hsl], 0ull, m);
                if (old == 0ull || old == m) gsl = (int)hsl;
                hsl = (hsl + 1) & (kJuncCap - 1);
              }
              s_jg[i] = gsl;
              if (gsl < 0) {
                atomicOr(err, kErrTopo);
              } else if constexpr (kElemLds) {
                atomicAdd(&gp[3 * gsl], s_jp[3 * i]);
                atomicAdd(&gp[3 * gsl + 1], s_jp[3 * i + 1]);
                atomicAdd(&gp[3 * gsl + 2], s_jp[3 * i + 2]);
                atomicAdd(&ga[gsl], s_ja[i]);
                atomicAdd(&gc[gsl], s_jc[i]);
              }
            }
            else {
              s_jg[i] = -1;
            }
            s_jm[i] = 0ull;
            if constexpr (kElemLds) {
              s_jc[i] = 0;
              s_ja[i] = 0.0;
              s_jp[3 * i] = s_jp[3 * i + 1] = s_jp[3 * i + 2] = 0ull;
            }
          }
          __syncthreads();
          // 4. with the convolution correction on: per cut cell the grain-table
          //    slot of its list and that list's first appearance; then every
          //    triangle's barycenter as a record (grain, slot), appended to the
          //    pool (one reservation per workgroup and pass)
          if (topo.corr) {
            int rslot[kCPT];
            int cnt = 0;
#pragma unroll
            for (int r = 0; r < kCPT; ++r) {
              // the grain-table slot of the cell's list, as the merge above left it
              rslot[r] = cjsl[r] >= 0 ? s_jg[cjsl[r]] : -1;
              if (rslot[r] < 0) continue;
              const int cell = (kl * m1 + (int)((code[r] >> 8) & 0xFF)) * m1 + (int)(code[r] >> 16);
              atomicMin(&topo.jord[(int64_t)g * kJuncCap + rslot[r]], cell);
              cnt += s_ntri[code[r] & 0xFFu];
            }
            int sc = cnt;  // inclusive scan over the wavefront
            for (int off = 1; off < 64; off <<= 1) {
              const int y = __shfl_up(sc, off);
              if (lane >= off) sc += y;
            }
            if (lane == 63) s_wave[wv] = sc;
            __syncthreads();
            int pos = sc - cnt, tot = 0;
            for (int w = 0; w < kHT / 64; ++w) {
              if (w < wv) pos += s_wave[w];
              tot += s_wave[w];
            }
            __shared__ unsigned long long s_rbase;
            if (tid == 0) s_rbase = tot ? atomicAdd(topo.raw_n, (unsigned long long)tot) : 0ull;
            __syncthreads();
            const unsigned long long rb = s_rbase;
            if (tot && rb + (unsigned long long)tot > (unsigned long long)topo.raw_cap) {
              if (tid == 0) atomicOr(err, kErrTri);
            } else if (tot) {
#pragma unroll
              for (int r = 0; r < kCPT; ++r) {
                if (rslot[r] < 0) continue;
                const unsigned ci = code[r] & 0xFFu;
                const int cyl = (code[r] >> 8) & 0xFF, cxl = (code[r] >> 16) & 0xFF;
                const int nt = s_ntri[ci];
                uint64_t row = s_tri[ci];
                for (int j = 0; j < nt; ++j, row >>= 12) {
                  double p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
                  edge_vertex<T, NMAX>((int)(row & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p0x, p0y, p0z);
                  edge_vertex<T, NMAX>((int)((row >> 4) & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p1x, p1y, p1z);
                  edge_vertex<T, NMAX>((int)((row >> 8) & 15u), L, cxl, cyl, n, x0 + cxl, y0 + cyl, z, p2x, p2y, p2z);
                  TriRec t;
                  t.x = (p0x + p1x + p2x) / 3;
                  t.y = (p0y + p1y + p2y) / 3;
                  t.z = (p0z + p1z + p2z) / 3;
                  t.g = g;
                  t.slot = rslot[r];
                  topo.raw[rb + pos++] = t;
                }
              }
            }
            __syncthreads();  // s_wave and s_rbase are reused
          }
          // 3. interfacial elements (src/grainHull.cpp:227-270) of the class
          //    without LDS for them: each cut cell adds its sums to its list's
          //    slot in the grain's junction table
#pragma unroll
          for (int r = 0; r < kCPT; ++r) {
            if constexpr (kElemLds) break;
            const int slot = cjsl[r] >= 0 ? s_jg[cjsl[r]] : -1;
            if (slot < 0) continue;
            double sx, sy, sz, sa;
            const int nt = cell_elems(code[r], sx, sy, sz, sa);
            atomicAdd(&gp[3 * slot], fixpos(sx));
            atomicAdd(&gp[3 * slot + 1], fixpos(sy));
            atomicAdd(&gp[3 * slot + 2], fixpos(sz));
            atomicAdd(&ga[slot], sa);
            atomicAdd(&gc[slot], nt);
          }
        }
        // exclusive scan of the per-lane triangle counts: wavefront, then workgroup
        int sc = mine;
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(sc, off);
          if (lane >= off) sc += y;
        }
        if (lane == 63) s_wave[wv] = sc;
        __syncthreads();
        int pos = sc - mine, nq = 0;
        for (int w = 0; w < kHT / 64; ++w) {
          if (w < wv) pos += s_wave[w];
          nq += s_wave[w];
        }
        if (nq) any = 1;
        // The triangles are listed cell by cell (entry = cell code | triangle
        // index << 24) in windows of kChunk entries; every lane then takes one
        // triangle at a time, so lanes do not wait on cells with more triangles.
        for (int w0 = 0; w0 < nq; w0 += kChunk) {
          int pp = pos;
#pragma unroll
          for (int r = 0; r < kCPT; ++r) {
            if (code[r] != 0u) {
              const int t = s_ntri[code[r] & 0xFFu];
              for (int j = 0; j < t; ++j, ++pp)
                if (pp >= w0 && pp < w0 + kChunk) q[pp - w0] = code[r] | (uint32_t)j << 24;
            }
          }
          __syncthreads();
          const int nw = min(kChunk, nq - w0);
          for (int w = tid; w < nw; w += kHT) {
            const uint32_t cd = q[w];
            const unsigned ci = cd & 0xFFu;
            const int yl = (cd >> 8) & 0xFF, xl = (cd >> 16) & 0xFF;
            const uint32_t row = (uint32_t)(s_tri[ci] >> (12 * (cd >> 24)));
            const int x = x0 + xl, y = y0 + yl;
            double p0x, p0y, p0z, p1x, p1y, p1z, p2x, p2y, p2z;
            edge_vertex<T, NMAX>((int)(row & 15u), L, xl, yl, n, x, y, z, p0x, p0y, p0z);
            edge_vertex<T, NMAX>((int)((row >> 4) & 15u), L, xl, yl, n, x, y, z, p1x, p1y, p1z);
            edge_vertex<T, NMAX>((int)((row >> 8) & 15u), L, xl, yl, n, x, y, z, p2x, p2y, p2z);
            const double v321 = p2x * p1y * p0z;
            const double v231 = p1x * p2y * p0z;
            const double v312 = p2x * p0y * p1z;
            const double v132 = p0x * p2y * p1z;
            const double v213 = p1x * p0y * p2z;
            const double v123 = p0x * p1y * p2z;
            vol.add((double)(1.0f / 6.0f) * (-v321 + v231 + v312 - v132 - v213 + v123));
            const double abx = p0x - p1x, aby = p0y - p1y, abz = p0z - p1z;
            const double acx = p0x - p2x, acy = p0y - p2y, acz = p0z - p2z;
            const double cx = aby * acz - abz * acy;
            const double cy = abz * acx - abx * acz;
            const double cz = abx * acy - aby * acx;
            area.add(sqrt(cx * cx + cy * cy + cz * cz) / 2.0);
          }
          __syncthreads();
        }
        __syncthreads();  // s_wave is read above before the next chunk writes it
      }
    }
    // wavefront then workgroup reduction in a fixed order
    double r4[4] = {vol.s, vol.c, area.s, area.c};
    for (int off = 32; off > 0; off >>= 1) {
      double s2 = __shfl_down(r4[0], off), c2 = __shfl_down(r4[1], off);
      double a2 = __shfl_down(r4[2], off), d2 = __shfl_down(r4[3], off);
      Acc va{r4[0], r4[1]}, aa{r4[2], r4[3]};
      va.merge(s2, c2);
      aa.merge(a2, d2);
      r4[0] = va.s; r4[1] = va.c; r4[2] = aa.s; r4[3] = aa.c;
    }
    if (lane == 0) {
      for (int qq = 0; qq < 4; ++qq) s_red[qq][wv] = r4[qq];
      s_any[wv] = any;
    }
    __syncthreads();
    if (tid == 0) {
      Acc va{s_red[0][0], s_red[1][0]}, aa{s_red[2][0], s_red[3][0]};
      int an = s_any[0];
      for (int w = 1; w < kHT / 64; ++w) {
        va.merge(s_red[0][w], s_red[1][w]);
        aa.merge(s_red[2][w], s_red[3][w]);
        an |= s_any[w];
      }
      double* pp = partials + 4 * slot;
      pp[0] = va.s; pp[1] = va.c; pp[2] = aa.s; pp[3] = aa.c;
      if (an) atomicOr(&exflag[g], 1);
    }
    __syncthreads();
  }
}

// Per grain: vanished -> deleted; alive -> take the comparison list as the new
// neighbour list, decide existence, finish volume/area, resize the output box
// to the set-comparison bounds (+-grid_blowup, clamped, src/box.cpp:857-877)
// made cubic.  mode selects the parts:
//   kFinBounds  the output box of every alive grain (it follows from the
//               set-comparison bounds alone, so it can precede the hull)
//   kFinState   deletion, list flip, existence, volume and area (after the hull)
//   kFinCoarsen extractContour as a grid coarsening calls it
//               (src/grainhdl.cpp:1481-1490) rather than level_set: no
//               deletion, the neighbour lists stay, and newb holds resizeGrid's
//               bounds, taken as they are
// With kFinState the output box is written for grains that keep existing only.
__global__ void k_finalize(SimParams sp, PlanDev P, DescDev din, int32_t* state, ListDev ld, int32_t* exflag,
                           const double* __restrict__ partials, const int32_t* __restrict__ newb, DescDev dout,
                           double* volume, double* area, int mode) {
  const bool coarsen = mode & kFinCoarsen, do_state = mode & kFinState, do_bounds = mode & kFinBounds;
  for (int g = 1 + blockIdx.x * blockDim.x + threadIdx.x; g <= sp.N; g += gridDim.x * blockDim.x) {
    const int s = state[g];
    if (s == kVanished) {
      if (do_state && !coarsen) state[g] = kDeleted;
      continue;
    }
    if (s == kDoomed) {
      if (do_state) {
        ld.sel[g] ^= 1;
        exflag[g] = 0;
        state[g] = kVanished;
      }
      continue;
    }
    if (s != kAlive) continue;
    if (do_state) {
      if (!coarsen) ld.sel[g] ^= 1;
      const int ex = exflag[g];
      exflag[g] = 0;
      if (!ex) {
        state[g] = kVanished;
        continue;
      }
      const int pos = P.grain_pos[g];
      int c = 0;
      while (c + 1 < kClasses && pos >= P.cls_start[c + 1]) ++c;
      const int idx = pos - P.cls_start[c];
      const int first = P.pre[kQLayers][P.cls_start[c] + c + idx];
      const int n = box_n(din.bounds, g);
      const double* pp = partials + 4 * (P.layer_base[c] + first);
      Acc va{pp[0], pp[1]}, aa{pp[2], pp[3]};
      for (int l = 1; l < n; ++l) {
        va.merge(pp[4 * l], pp[4 * l + 1]);
        aa.merge(pp[4 * l + 2], pp[4 * l + 3]);
      }
      const double h = sp.h;
      volume[g] = fabs(va.s + va.c) * h * h * h;
      area[g] = (aa.s + aa.c) * h * h;
    }
    if (!do_bounds) continue;
    const int32_t* nb = newb + 6 * g;
    // Before the hull: a grain without a voxel >= 0 (empty set-comparison
    // bounds) has no hull and vanishes; it is left out of the output plan.
    if (!do_state && (nb[0] > nb[3] || nb[1] > nb[4] || nb[2] > nb[5])) {
      state[g] = kDoomed;
      continue;
    }
    int bb[6] = {nb[0] - sp.gb, nb[1] - sp.gb, nb[2] - sp.gb, nb[3] + sp.gb, nb[4] + sp.gb, nb[5] + sp.gb};
    if (coarsen) {
      for (int a = 0; a < 6; ++a) bb[a] = nb[a];
    } else {
      for (int a = 0; a < 3; ++a) {
        if (bb[a] < 0) bb[a] = 0;
        if (bb[a + 3] >= sp.ngp) bb[a + 3] = sp.ngp - 1;
      }
    }
    cube_bounds(bb, sp.ngp);
    for (int a = 0; a < 6; ++a) dout.bounds[6 * g + a] = bb[a];
  }
}

}  // namespace

void upload_mc_tables() {
  uint64_t tri[256];
  for (int c = 0; c < 256; ++c) {
    uint64_t row = ~0ull;
    const char* t = kTriCases[c];
    for (int i = 0; t[i]; ++i) {
      const char ch = t[i];
      const uint64_t e = (uint64_t)((ch >= '0' && ch <= '9') ? ch - '0' : ch - 'a' + 10);
      row = (row & ~(15ull << (4 * i))) | e << (4 * i);
    }
    tri[c] = row;
  }
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_tri), tri, sizeof(tri));
}

template <class T, bool TOPO>
void hull_classes(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din, int32_t* exflag,
                  double* partials, const IdlDev& il, const TopoDev& topo, int32_t* err, int c0, int c1) {
  if (c0 <= 0 && 0 < c1)
    hipLaunchKernelGGL((k_hull<T, 32, TOPO>), dim3(tile_grid(L, kGridPerCU, 0)), dim3(kHT), 0, L.stream, sp, P, 0,
                       din, exflag, partials, il, topo, err);
  if (c0 <= 1 && 1 < c1)
    hipLaunchKernelGGL((k_hull<T, 48, TOPO>), dim3(tile_grid(L, kGridPerCU, 1)), dim3(kHT), 0, L.stream, sp, P, 1,
                       din, exflag, partials, il, topo, err);
  if (c0 <= 2 && 2 < c1)
    hipLaunchKernelGGL((k_hull<T, 64, TOPO>), dim3(tile_grid(L, kGridPerCU, 2)), dim3(kHT), 0, L.stream, sp, P, 2,
                       din, exflag, partials, il, topo, err);
  if (c0 <= 3 && 3 < c1)
    hipLaunchKernelGGL((k_hull<T, 96, TOPO>), dim3(tile_grid(L, kGridPerCU, 3)), dim3(kHT), 0, L.stream, sp, P, 3,
                       din, exflag, partials, il, topo, err);
}

// Clears the grains' topology tables before a hull with topology on.
__global__ __launch_bounds__(256) void k_topo_clear(SimParams sp, TopoDev topo) {
  const int64_t n = (int64_t)(sp.N + 1) * kTopoCap, nj = (int64_t)(sp.N + 1) * kJuncCap;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nj; i += (int64_t)gridDim.x * blockDim.x) {
    if (i < n) {
      topo.key[i] = -1;
      topo.ord[i] = INT_MAX;
    }
    topo.jkey[i] = 0ull;
    topo.jcnt[i] = 0;
    if (topo.corr) topo.jord[i] = INT_MAX;
    if (topo.corr && i == 0) *topo.raw_n = 0ull;
    topo.jpos[3 * i] = 0ull;
    topo.jpos[3 * i + 1] = 0ull;
    topo.jpos[3 * i + 2] = 0ull;
    topo.jarea[i] = 0.0;
  }
}

void launch_topo_clear(const Launch& L, const SimParams& sp, const TopoDev& topo) {
  const int64_t n = (int64_t)(sp.N + 1) * kJuncCap;
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_topo_clear, dim3(blocks), dim3(256), 0, L.stream, sp, topo);
}

void launch_hull(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din, int32_t* exflag,
                 double* partials, const IdlDev& il, const TopoDev& topo, int32_t* err, bool fp32, int c0, int c1) {
  if (fp32) {
    if (topo.on) hull_classes<float, true>(L, sp, P, din, exflag, partials, il, topo, err, c0, c1);
    else hull_classes<float, false>(L, sp, P, din, exflag, partials, il, topo, err, c0, c1);
  } else {
    if (topo.on) hull_classes<double, true>(L, sp, P, din, exflag, partials, il, topo, err, c0, c1);
    else hull_classes<double, false>(L, sp, P, din, exflag, partials, il, topo, err, c0, c1);
  }
}

void launch_finalize(const Launch& L, const SimParams& sp, const PlanDev& P, int32_t* state, const ListDev& ld,
                     int32_t* exflag, const double* partials, const int32_t* newb, const DescDev& din,
                     const DescDev& dout, double* volume, double* area, int mode) {
  int blocks = (sp.N + 255) / 256;
  hipLaunchKernelGGL(k_finalize, dim3(blocks), dim3(256), 0, L.stream, sp, P, din, state, ld, exflag, partials,
                     newb, dout, volume, area, mode);
}

}  // namespace gg
