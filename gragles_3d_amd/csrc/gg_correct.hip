// Convolution correction: the correction branch of LSbox::executeConvolution
// (src/box.cpp:428-536) on the interfacial elements of every grain's last hull
// (GrainHull::computeGrainBoundaryElements / subDivideTrianglesToInterfacialElements
// / computeJunctionPosition with mergeJunction, src/grainHull.cpp:187-274,
// 483-531; element constructors src/InterfacialElement.cpp:52-404) found with
// GrainHull::projectPointToGrainBoundary (src/grainHull.cpp:570-660).
//
// After each hull that ran with the correction on (level set, grid coarsening):
//   k_elements     one workgroup per grain: the hull's distinct cut-cell id
//                  lists (the grain's junction table, with each list's first
//                  appearance) in first-appearance order become elements by
//                  size (2 grains: boundary, 3: triple line, 4: quadruple
//                  junction, else high-order junction) with their energy and
//                  mobility; quadruple junctions closer than 2 voxels merge as
//                  mergeJunction merges them; elements are laid out in the order
//                  the projection visits them (high-order junctions, quadruple
//                  junctions, triple lines, boundaries) with the start of each
//                  element's barycenters, and the grain reserves its range of
//                  the barycenter array.
//   k_tri_scatter  every triangle record of the hull goes to its element's
//                  range (the order inside an element does not matter: all its
//                  barycenters carry the element's weights).
// After each convolution:
//   k_correct      one workgroup per (grain, z-layer) of the output box: the
//                  band voxels (|phi_in| < delta) are queued with the slot of
//                  their IDLocal id in the grain's hull-id table and grouped by
//                  slot (counting sort in LDS), so a wavefront's 64 voxels
//                  mostly border the same grain; each lane takes one, walks
//                  the grain's elements in projection order (wave-uniform
//                  loop that skips the elements no lane needs; the
//                  barycenters are read with wave-uniform loads, four per
//                  round) and keeps the first nearest barycenter (squared
//                  distance, strict '<') among the elements that border its
//                  id; then phi_out = phi_in +
//                  (phi_out - phi_in) * w_energy * w_mobility and, with stored
//                  energies on, minus (SEE_self - SEE_id) * w_mobility * dt /
//                  TimeSlope^2.  The weights are GBInfo's as get_GBInfo hands
//                  them over: w_energy = the element's mobility, w_mobility =
//                  its energy (src/InterfacialElement.h:22-31, 70-72).
//
// FP64 arithmetic throughout, in the reference's operation order; the junction
// positions (used only for the merge test) are the hull's barycenter sums
// (fixed point, summed in any order to the same value) over their count.
#include "gg_internal.h"

namespace gg {
namespace {

constexpr int kEW = 256;                           // threads of an element-build workgroup
constexpr int kRec = kJuncCap + kJuncCap / 2;      // lists + high-order junctions made by merging
constexpr int kTopoCapBitsC = 6;
static_assert((1 << kTopoCapBitsC) == kTopoCap, "grain topology table size");
constexpr int kCorrGridPerCU = 512;  // workgroups per CU of a correction launch (512 against 128: stored-energy step 202.0 -> 199.5 ms, timed on one box)
constexpr int kCellBits = 14;                      // a layer cell index (n^2 <= 9216) in a queue entry
static_assert(kMaxN * kMaxN < (1 << kCellBits), "queue entry layout");

enum { kClsGB = 0, kClsTL = 1, kClsQJ = 2, kClsHJ = 3 };

__device__ inline double pair_gam(const CorrDev& cd, int N, unsigned a, unsigned b) {
  return cd.aniso ? cd.gam[(int64_t)a * (N + 1) + b] : 1.0;
}
__device__ inline double pair_mob(const CorrDev& cd, int N, unsigned a, unsigned b) {
  return cd.aniso ? cd.mob[(int64_t)a * (N + 1) + b] : 1.0;
}

// TripleLine::computeMobility / computeEnergy (src/InterfacialElement.cpp:118-164)
__device__ inline void triple_line(const CorrDev& cd, int N, unsigned me, unsigned n0, unsigned n1, double& energy,
                                   double& mobility) {
  double avg = 0;
  avg += pair_mob(cd, N, me, n0);
  avg += pair_mob(cd, N, n0, n1);
  avg += pair_mob(cd, N, me, n1);
  avg /= 3;
  mobility = 1. / (cd.tl_inv + 1. / avg);
  const double g0 = pair_gam(cd, N, me, n0), g1 = pair_gam(cd, N, n0, n1), g2 = pair_gam(cd, N, me, n1);
  double sigma = g0 - g1 + g2;
  if (sigma < 0.01) sigma = 0.01;
  if (sigma > 1.0) sigma = 1.0;
  energy = sigma;
}

__global__ __launch_bounds__(kEW) void k_elements(SimParams sp, const int32_t* __restrict__ state, TopoDev topo,
                                                  CorrDev cd, int32_t* err) {
  __shared__ int s_key[kTopoCap];
  __shared__ int s_nl, s_self, s_E;
  __shared__ short s_ls[kJuncCap];      // junction-table slot of each list (compacted)
  __shared__ int s_lo[kJuncCap];        // its first appearance
  __shared__ short s_sorted[kJuncCap];  // compacted index of the r-th list in first-appearance order
  // records: the lists in first-appearance order (index r < nl), then the
  // high-order junctions made by merging
  __shared__ unsigned long long s_em[kRec];  // neighbour set (slots of the grain's hull-id table)
  __shared__ double s_en[kRec], s_mo[kRec];
  __shared__ double s_sx[kRec], s_sy[kRec], s_sz[kRec];  // barycenter sums (junctions)
  __shared__ int s_cnt[kRec];
  __shared__ signed char s_cls[kRec];
  __shared__ short s_red[kRec];   // record it was merged into (-1: none)
  __shared__ short s_elem[kRec];  // element index of a record that is an element
  __shared__ short s_hj[kRec], s_qj[kJuncCap];
  __shared__ short s_ord[kJuncCap];  // element -> record
  __shared__ int s_efp[kJuncCap + 1];  // first barycenter of each element (whole chunks)
  __shared__ int64_t s_tb;
  const int N = sp.N;
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid == 0) *cd.ready = 1;
  for (int g = 1 + blockIdx.x; g <= N; g += gridDim.x) {
    if (state[g] != kAlive) {
      if (tid == 0) cd.ecount[g] = 0;
      continue;
    }
    const int64_t jb = (int64_t)g * kJuncCap;
    if (tid < kTopoCap) s_key[tid] = topo.key[(int64_t)g * kTopoCap + tid];
    if (tid == 0) {
      s_nl = 0;
      s_self = -1;
    }
    __syncthreads();
    if (tid < kTopoCap && s_key[tid] == g) s_self = tid;
    for (int i = tid; i < kJuncCap; i += kEW) {
      if (topo.jkey[jb + i] != 0ull) {
        const int q = atomicAdd(&s_nl, 1);
        s_ls[q] = (short)i;
        s_lo[q] = topo.jord[jb + i];
      }
    }
    __syncthreads();
    const int nl = s_nl;
    for (int q = tid; q < nl; q += kEW) {
      const int o = s_lo[q];
      int r = 0;
      for (int j = 0; j < nl; ++j) r += (s_lo[j] < o) | ((s_lo[j] == o) & (j < q));
      s_sorted[r] = (short)q;
    }
    __syncthreads();
    // each list as an element: class by its number of grains, neighbour ids
    // ascending, energy and mobility (src/InterfacialElement.cpp:52-404)
    const unsigned me = (unsigned)g;
    for (int r = tid; r < nl; r += kEW) {
      const int slot = s_ls[s_sorted[r]];
      const unsigned long long m = topo.jkey[jb + slot];
      const int self = s_self;
      const unsigned long long nbm = self >= 0 ? m & ~(1ull << self) : m;
      const int size = __popcll(m);
      unsigned o[3] = {0u, 0u, 0u};
      int no = 0;
      if (size >= 2 && size <= 4) {
        for (unsigned long long t = nbm; t && no < 3; t &= t - 1) {
          unsigned v = (unsigned)s_key[__builtin_ctzll(t)];
          int p = no++;
          while (p > 0 && o[p - 1] > v) {
            o[p] = o[p - 1];
            --p;
          }
          o[p] = v;
        }
      }
      int cls;
      double en = 1.0, mo = 1.0;
      if (size == 2 && no == 1) {  // GrainBoundary
        cls = kClsGB;
        mo = pair_mob(cd, N, me, o[0]);
        en = pair_gam(cd, N, me, o[0]);
      } else if (size == 3 && no == 2) {  // TripleLine
        cls = kClsTL;
        triple_line(cd, N, me, o[0], o[1], en, mo);
      } else if (size == 4 && no == 3) {  // QuadrupleJunction: means over its triple lines
        cls = kClsQJ;
        double e1, m1, e2, m2, e3, m3;
        triple_line(cd, N, me, o[0], o[1], e1, m1);
        triple_line(cd, N, me, o[1], o[2], e2, m2);
        triple_line(cd, N, me, o[0], o[2], e3, m3);
        mo = (m1 + m2 + m3) / 3.0;
        en = (e1 + e2 + e3) / 3.0;
      } else {  // HighOrderJunction (other sizes)
        cls = kClsHJ;
      }
      s_em[r] = nbm;
      s_cls[r] = (signed char)cls;
      s_en[r] = en;
      s_mo[r] = mo;
      s_cnt[r] = topo.jcnt[jb + slot];
      s_sx[r] = (double)(long long)topo.jpos[3 * (jb + slot)] / kPosScale;
      s_sy[r] = (double)(long long)topo.jpos[3 * (jb + slot) + 1] / kPosScale;
      s_sz[r] = (double)(long long)topo.jpos[3 * (jb + slot) + 2] / kPosScale;
      s_red[r] = -1;
    }
    __syncthreads();
    if (tid == 0) {
      int nh = 0, nq = 0, nrec = nl;
      for (int r = 0; r < nl; ++r) {
        if (s_cls[r] == kClsHJ) s_hj[nh++] = (short)r;
        else if (s_cls[r] == kClsQJ) s_qj[nq++] = (short)r;
      }
      // junction position = mean barycenter (computePosition); distance of two
      auto near = [&](int a, int b) {
        const double na = (double)s_cnt[a], nb = (double)s_cnt[b];
        const double dx = s_sx[a] / na - s_sx[b] / nb, dy = s_sy[a] / na - s_sy[b] / nb,
                     dz = s_sz[a] / na - s_sz[b] / nb;
        return sqrt(dx * dx + dy * dy + dz * dz) < 2;
      };
      auto absorb = [&](int h, int b) {  // mergeWith: b's barycenters and ids join h
        s_cnt[h] += s_cnt[b];
        s_sx[h] += s_sx[b];
        s_sy[h] += s_sy[b];
        s_sz[h] += s_sz[b];
        s_em[h] |= s_em[b];
        s_en[h] = s_mo[h] = 1.0;
        s_red[b] = (short)h;
      };
      // mergeJunction (src/grainHull.cpp:483-531)
      for (int i = 0; i < nq; i++) {
        for (int j = i + 1; j < nq; j++) {
          const int A = s_qj[i], B = s_qj[j];
          if (near(A, B)) {
            // HighOrderJunction(A, B): A's ids (B adds only ids A has), A's then
            // B's barycenters
            const int h = nrec++;
            s_em[h] = s_em[A];
            s_cnt[h] = s_cnt[A];
            s_sx[h] = s_sx[A];
            s_sy[h] = s_sy[A];
            s_sz[h] = s_sz[A];
            s_cls[h] = kClsHJ;
            s_red[h] = -1;
            absorb(h, B);
            s_em[h] = s_em[A];
            s_red[A] = (short)h;
            for (int t = j; t + 1 < nq; ++t) s_qj[t] = s_qj[t + 1];
            --nq;
            for (int t = i; t + 1 < nq; ++t) s_qj[t] = s_qj[t + 1];
            --nq;
            i--;
            s_hj[nh++] = (short)h;
            break;
          }
        }
      }
      for (int k = 0; k < nh; k++)
        for (int i = 0; i < nq; i++)
          if (near(s_qj[i], s_hj[k])) {
            absorb(s_hj[k], s_qj[i]);
            for (int t = i; t + 1 < nq; ++t) s_qj[t] = s_qj[t + 1];
            --nq;
            i--;
          }
      for (int i = 0; i < nh; i++)
        for (int k = i + 1; k < nh; k++)
          if (near(s_hj[i], s_hj[k])) {
            absorb(s_hj[i], s_hj[k]);
            for (int t = k; t + 1 < nh; ++t) s_hj[t] = s_hj[t + 1];
            --nh;
            k--;
          }
      // projection order: high-order junctions, quadruple junctions, triple
      // lines, boundaries
      int E = 0;
      for (int t = 0; t < nh; ++t) s_ord[E++] = s_hj[t];
      for (int t = 0; t < nq; ++t) s_ord[E++] = s_qj[t];
      for (int r = 0; r < nl; ++r)
        if (s_cls[r] == kClsTL) s_ord[E++] = (short)r;
      const int ntl = E - nh - nq;
      for (int r = 0; r < nl; ++r)
        if (s_cls[r] == kClsGB) s_ord[E++] = (short)r;
      int64_t acc = 0;
      int32_t* ef = cd.efirst + (int64_t)g * (kElemCap + 1);
      for (int e = 0; e < E; ++e) {
        s_elem[s_ord[e]] = (short)e;
        ef[e] = (int32_t)acc;
        s_efp[e] = (int32_t)acc;
        acc += (s_cnt[s_ord[e]] + kChunk - 1) / kChunk * kChunk;  // whole chunks
      }
      ef[E] = (int32_t)acc;
      s_efp[E] = (int32_t)acc;
      const unsigned long long tb = atomicAdd(cd.tri_n, (unsigned long long)acc);
      if (tb + (unsigned long long)acc > (unsigned long long)cd.tri_cap) atomicOr(err, kErrTri);
      cd.tbase[g] = (int64_t)tb;
      s_tb = (int64_t)tb;
      cd.ecount[g] = E;
      cd.eclass[4 * g] = nh;
      cd.eclass[4 * g + 1] = nq;
      cd.eclass[4 * g + 2] = ntl;
      s_E = E;
    }
    __syncthreads();
    const int E = s_E;
    const int64_t eb = (int64_t)g * kElemCap;
    for (int e = tid; e < E; e += kEW) {
      const int rec = s_ord[e];
      cd.emask[eb + e] = s_em[rec];
      cd.energy[eb + e] = s_en[rec];
      cd.mobility[eb + e] = s_mo[rec];
      cd.ecur[eb + e] = 0;
      const double n = (double)s_cnt[rec];
      const bool junction = s_cls[rec] >= kClsQJ;
      cd.epos[3 * (eb + e)] = junction ? s_sx[rec] / n : 0.0;
      cd.epos[3 * (eb + e) + 1] = junction ? s_sy[rec] / n : 0.0;
      cd.epos[3 * (eb + e) + 2] = junction ? s_sz[rec] / n : 0.0;
    }
    for (int r = tid; r < nl; r += kEW) {
      int rec = r;
      while (s_red[rec] >= 0) rec = s_red[rec];
      cd.slot_elem[jb + s_ls[s_sorted[r]]] = s_elem[rec];
    }
    // the padding points that complete each element's last chunk
    for (int e = tid; e < E; e += kEW) {
      const int64_t a = s_tb + s_efp[e] + s_cnt[s_ord[e]], b = s_tb + s_efp[e + 1];
      for (int64_t q = a; q < b && q < cd.tri_cap; ++q) {
        cd.tri[3 * q] = kFarPoint;
        cd.tri[3 * q + 1] = kFarPoint;
        cd.tri[3 * q + 2] = kFarPoint;
      }
    }
    __syncthreads();  // the LDS tables are refilled for the next grain
  }
}

// The barycenter pool's fill counter (a kernel, so that a captured step graph
// holds no memset node).
__global__ void k_clear_counter(unsigned long long* c) {
  if (threadIdx.x == 0) *c = 0ull;
}

// Every triangle record of the last hull into its element's barycenter range.
// The records of one hull tile are consecutive in the pool and in cell order;
// a wavefront reserves one run per (grain, element) among its 64 records, so
// an element's range is made of such runs in the hull's order and consecutive
// barycenters lie close together (which keeps the chunk boxes small).
__global__ __launch_bounds__(256) void k_tri_scatter(TopoDev topo, CorrDev cd) {
  unsigned long long n = *topo.raw_n;
  if (n > (unsigned long long)topo.raw_cap) n = (unsigned long long)topo.raw_cap;
  const int lane = threadIdx.x & 63;
  for (unsigned long long base = blockIdx.x * 256ull; base < n; base += (unsigned long long)gridDim.x * 256ull) {
    const unsigned long long r = base + threadIdx.x;
    bool valid = r < n;
    TriRec t{};
    int e = -1;
    unsigned key = 0u;
    if (valid) {
      t = topo.raw[r];
      e = cd.slot_elem[(int64_t)t.g * kJuncCap + t.slot];
      valid = (unsigned)e < (unsigned)cd.ecount[t.g];  // false only after a table overflow (device error set)
      key = (unsigned)t.g * (unsigned)kElemCap + (unsigned)e;
    }
    int64_t pos = 0;
    bool done = !valid;
    for (;;) {  // one round per distinct (grain, element) among the wavefront's records
      const unsigned long long pend = __ballot(!done);
      if (pend == 0ull) break;
      const int leader = __ffsll((long long)pend) - 1;
      const unsigned k0 = __shfl(key, leader);
      const bool mine = !done && key == k0;
      const unsigned long long m = __ballot(mine);
      int b0 = 0;
      if (lane == leader) b0 = atomicAdd(&cd.ecur[k0], __popcll(m));
      b0 = __shfl(b0, leader);
      if (mine) {
        pos = b0 + __popcll(m & ((1ull << lane) - 1ull));
        done = true;
      }
    }
    if (valid) {
      const int64_t q = cd.tbase[t.g] + cd.efirst[(int64_t)t.g * (kElemCap + 1) + e] + pos;
      if (q < cd.tri_cap) {
        cd.tri[3 * q] = t.x;
        cd.tri[3 * q + 1] = t.y;
        cd.tri[3 * q + 2] = t.z;
      }
    }
  }
}

// Bounding box of every chunk of kChunk barycenters (padding points left
// out), widened by 1e-3 voxel so the FP32 box holds its FP64 points.
__global__ __launch_bounds__(256) void k_chunk_boxes(CorrDev cd) {
  unsigned long long nt = *cd.tri_n;
  if (nt > (unsigned long long)cd.tri_cap) nt = (unsigned long long)cd.tri_cap;
  const int64_t nc = (int64_t)(nt / kChunk);
  for (int64_t c = blockIdx.x * 256ll + threadIdx.x; c < nc; c += (int64_t)gridDim.x * 256ll) {
    float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
    const double* q = cd.tri + 3 * c * kChunk;
    for (int j = 0; j < kChunk; ++j) {
      if (q[3 * j] >= 0.5 * kFarPoint) continue;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float v = (float)q[3 * j + a];
        lo[a] = fminf(lo[a], v - 1e-3f);
        hi[a] = fmaxf(hi[a], v + 1e-3f);
      }
    }
    float* bx = cd.cbox + 6 * c;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      bx[a] = lo[a];
      bx[3 + a] = hi[a];
    }
  }
}

// Queue entry of a band voxel: its layer cell, and its IDLocal id's slot in the
// grain's hull-id table + 1 (0: not there).  Bucket of the slot grouping: the
// slot, or kTopoCap for an id the hull does not name.
__device__ inline int bucket_of(uint32_t qv) {
  const int f = (int)(qv >> kCellBits);
  return f == 0 ? kTopoCap : f - 1;
}

template <class T, int NMAX>
__global__ __launch_bounds__(256) void k_correct(SimParams sp, PlanDev P, int cls, const int32_t* __restrict__ state,
                                                 DescDev din, DescDev dout, IdlDev il, TopoDev topo, CorrDev cd) {
  if (*cd.ready == 0) return;  // no hull yet (run_sim's first step)
  constexpr int kBuckets = kTopoCap + 1;  // a slot, or the last bucket: an id the hull does not name
  static_assert(kTopoCap == 64, "one wavefront scans the slot buckets");
  __shared__ int s_key[kTopoCap];
  __shared__ unsigned long long s_em[kElemCap];
  __shared__ double s_en[kElemCap], s_mo[kElemCap];
  __shared__ int s_ef[kElemCap + 1];
  __shared__ uint32_t s_raw[NMAX * NMAX];  // band voxels in layer order
  __shared__ uint32_t s_q[NMAX * NMAX];    // the same, grouped by slot
  __shared__ int s_cnt[kBuckets], s_at[kBuckets];
  __shared__ int s_nq;
  const int total = P.totals[kQLayers * kClasses + cls];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const double delta = sp.delta;
  // consecutive tiles (the layers of one grain) on one XCD: the layer two
  // tiles share, and the grain's tables, come from that XCD's L2
  const TileRange tr = xcd_tiles(total);
  for (int tile = tr.first; tile < tr.end; tile += tr.step) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* b = dout.bounds + 6 * g;
    const int x0 = b[0], y0 = b[1], z0 = b[2], n = b[3] - b[0], nn2 = n * n;
    // a grain whose barycenters did not fit the pool (the device error is set)
    // is corrected with the default weights
    const int64_t tb = cd.tbase[g];
    int E = cd.ecount[g];
    if (tb + cd.efirst[(int64_t)g * (kElemCap + 1) + E] > cd.tri_cap) E = 0;
    for (int i = tid; i < kTopoCap; i += 256) s_key[i] = topo.key[(int64_t)g * kTopoCap + i];
    for (int e = tid; e < E; e += 256) {
      s_em[e] = cd.emask[(int64_t)g * kElemCap + e];
      s_en[e] = cd.energy[(int64_t)g * kElemCap + e];
      s_mo[e] = cd.mobility[(int64_t)g * kElemCap + e];
    }
    for (int e = tid; e <= E; e += 256) s_ef[e] = cd.efirst[(int64_t)g * (kElemCap + 1) + e];
    if (tid < kBuckets) s_cnt[tid] = 0;
    if (tid == 0) s_nq = 0;
    __syncthreads();
    // IDLocal as the resize to the output bounds leaves it: the old content at
    // the same linear index, 0 past its old size
    const int32_t* ob = il.bounds + 6 * g;
    const int64_t oldsize = (int64_t)(ob[3] - ob[0]) * (ob[4] - ob[1]) * (ob[5] - ob[2]);
    const uint32_t* __restrict__ idl = il.arena + il.off[g];
    const T* __restrict__ in = ain + din.off[g];
    T* __restrict__ out = aout + dout.off[g];
    const int64_t lbase = (int64_t)kl * nn2;
    for (int c = tid; c < nn2; c += 256) {
      const double val = (double)in[lbase + c];
      if (fabs(val) < delta) {
        const int64_t li = lbase + c;
        const uint32_t id = li < oldsize ? idl[li] : 0u;
        int slot = -1;
        uint32_t hsl = (id * 2654435761u) >> (32 - kTopoCapBitsC);
        for (int pr = 0; pr < kTopoCap; ++pr) {
          const int k = s_key[hsl];
          if (k == (int)id) {
            slot = (int)hsl;
            break;
          }
          if (k == -1) break;
          hsl = (hsl + 1) & (kTopoCap - 1);
        }
        const uint32_t qv = (uint32_t)c | (uint32_t)(slot + 1) << kCellBits;
        atomicAdd(&s_cnt[bucket_of(qv)], 1);
        s_raw[atomicAdd(&s_nq, 1)] = qv;
      }
    }
    __syncthreads();
    // group the queue by slot (counting sort), so that a wavefront's voxels
    // mostly border one grain and it walks only that grain's elements
    if (wv == 0) {
      const int v = s_cnt[lane];
      int incl = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(incl, d, 64);
        if (lane >= d) incl += u;
      }
      s_at[lane] = incl - v;
      if (lane == 63) s_at[kTopoCap] = incl;
    }
    __syncthreads();
    const int nq = s_nq;
    for (int w = tid; w < nq; w += 256) {
      const uint32_t qv = s_raw[w];
      s_q[atomicAdd(&s_at[bucket_of(qv)], 1)] = qv;
    }
    __syncthreads();
    const double* __restrict__ tri = cd.tri + 3 * tb;
    const double see_g = cd.see_on ? cd.see[g] : 0.0;
    for (int w0 = wv * 64; w0 < nq; w0 += 256) {
      const int w = w0 + lane;  // a wavefront takes 64 consecutive queue entries
      const bool act = w < nq;
      const uint32_t qv = act ? s_q[w] : 0u;
      const int c = (int)(qv & ((1u << kCellBits) - 1u));
      const int slot = (int)(qv >> kCellBits) - 1;
      const int yl = c / n, xl = c - yl * n;
      // Vector3d(i, j, k): row, column, depth, against barycenters (x, y, z)
      const double px = (double)(y0 + yl), py = (double)(x0 + xl), pz = (double)(z0 + kl);
      // the first nearest barycenter in the projection order is the smallest
      // (squared distance, element) pair, found here in any order: elements
      // from boundaries back to junctions, skipping every chunk whose box no
      // lane can reach below its bound or its current pair
      double dmin = 10000000.0;
      int emin = -1;
      const unsigned long long bit = (act && slot >= 0) ? 1ull << slot : 0ull;
      const float* __restrict__ cbx = cd.cbox + 6 * (tb / kChunk);
      // A chunk whose FP32 box lies beyond the lane's current pair cannot hold
      // the answer.  Box distances only decide which chunks to skip, so each is
      // taken with a margin that covers its rounding (about 3 ulp): the skip is
      // conservative, and every barycenter that can win is compared in FP64.
      // (A first pass that bounded the search by the nearest box's farthest
      // corner cost more than it saved.)
      const float fpx = (float)px, fpy = (float)py, fpz = (float)pz;  // integers: exact
      for (int e = E - 1; e >= 0; --e) {
        const bool need = (s_em[e] & bit) != 0ull;
        if (__ballot(need) == 0ull) continue;
        const int ch1 = __builtin_amdgcn_readfirstlane(s_ef[e + 1]) / kChunk;
        for (int ch = __builtin_amdgcn_readfirstlane(s_ef[e]) / kChunk; ch < ch1; ++ch) {
          const float* bx = cbx + 6 * ch;
          const float ex = fmaxf(fmaxf(bx[0] - fpx, fpx - bx[3]), 0.0f);
          const float ey = fmaxf(fmaxf(bx[1] - fpy, fpy - bx[4]), 0.0f);
          const float ez = fmaxf(fmaxf(bx[2] - fpz, fpz - bx[5]), 0.0f);
          const float lb = (ex * ex + ey * ey + ez * ez) * (1.0f - 1e-6f) - 1e-6f;  // below the true one
          const bool want = need && (double)lb <= dmin;
          if (__ballot(want) == 0ull) continue;
          const double* q = tri + 3 * kChunk * ch;
#pragma unroll
          for (int u0 = 0; u0 < kChunk; u0 += 4) {
            double d[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const double dx = px - q[3 * (u0 + u)], dy = py - q[3 * (u0 + u) + 1], dz = pz - q[3 * (u0 + u) + 2];
              d[u] = dx * dx + dy * dy + dz * dz;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              if (want && (d[u] < dmin || (d[u] == dmin && e < emin))) {
                dmin = d[u];
                emin = e;
              }
          }
        }
      }
      const double w_energy = emin >= 0 ? s_mo[emin] : 1.0, w_mobility = emin >= 0 ? s_en[emin] : 1.0;
      if (act) {
        const int64_t li = lbase + c;
        const double val = (double)in[li];
        double o = (double)out[li];
        o = val + (o - val) * w_energy * w_mobility;
        if (cd.see_on) {
          const uint32_t id = li < oldsize ? idl[li] : 0u;
          const bool nb = id == 0u || (id <= (uint32_t)sp.N && state[id] != kDeleted);
          double f = 0;
          if (nb) f = (see_g - cd.see[id]) * w_mobility * cd.dt / cd.tsq;
          o = o - f;
        }
        out[li] = (T)o;
      }
    }
    __syncthreads();  // the queues and the element tables are refilled for the next tile
  }
}

template <class T>
void correct_classes(const Launch& L, const SimParams& sp, const PlanDev& P, const int32_t* state, const DescDev& din,
                     const DescDev& dout, const IdlDev& il, const TopoDev& topo, const CorrDev& cd) {
  hipLaunchKernelGGL((k_correct<T, 32>), dim3(tile_grid(L, kCorrGridPerCU, 0)), dim3(256), 0, L.stream, sp, P, 0,
                     state, din, dout, il, topo, cd);
  hipLaunchKernelGGL((k_correct<T, 48>), dim3(tile_grid(L, kCorrGridPerCU, 1)), dim3(256), 0, L.stream, sp, P, 1,
                     state, din, dout, il, topo, cd);
  hipLaunchKernelGGL((k_correct<T, 64>), dim3(tile_grid(L, kCorrGridPerCU, 2)), dim3(256), 0, L.stream, sp, P, 2,
                     state, din, dout, il, topo, cd);
  hipLaunchKernelGGL((k_correct<T, 96>), dim3(tile_grid(L, kCorrGridPerCU, 3)), dim3(256), 0, L.stream, sp, P, 3,
                     state, din, dout, il, topo, cd);
}

}  // namespace

void launch_elements(const Launch& L, const SimParams& sp, const int32_t* state, const TopoDev& topo,
                     const CorrDev& cd, int32_t* err) {
  hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(64), 0, L.stream, cd.tri_n);
  const int blocks = sp.N < L.grid ? sp.N : L.grid;
  hipLaunchKernelGGL(k_elements, dim3(blocks > 0 ? blocks : 1), dim3(kEW), 0, L.stream, sp, state, topo, cd, err);
  hipLaunchKernelGGL(k_tri_scatter, dim3(L.grid), dim3(256), 0, L.stream, topo, cd);
  hipLaunchKernelGGL(k_chunk_boxes, dim3(L.grid), dim3(256), 0, L.stream, cd);
}

void launch_correction(const Launch& L, const SimParams& sp, const PlanDev& P, const int32_t* state,
                       const DescDev& din, const DescDev& dout, const IdlDev& il_stale, const TopoDev& topo,
                       const CorrDev& cd, bool fp32) {
  if (fp32) correct_classes<float>(L, sp, P, state, din, dout, il_stale, topo, cd);
  else correct_classes<double>(L, sp, P, state, din, dout, il_stale, topo, cd);
}

}  // namespace gg
