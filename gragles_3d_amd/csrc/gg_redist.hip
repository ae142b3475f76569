// Redistancing of every grain box (LSbox::executeRedistancing,
// src/box.cpp:2322-2586).
//
// The reference runs, per z-layer k in [izmin, izmax-1), an x-forward, an
// x-backward, a y-forward and a y-backward sweep (each line independent), and
// then a z-forward and a z-backward sweep over all layers (each (y,x) column
// independent), and finally clamps to +-delta.  Layers are independent in the
// first phase and columns in the second, so:
//   k_redist_xy : one workgroup per (grain, output z-layer); the output layer
//                 and the in/out-intersection of the input layer sit in LDS;
//                 one lane per row for the x sweeps, one lane per column for
//                 the y sweeps, each line swept sequentially in the reference's
//                 order (so FP64 results are bit-identical).
//   k_redist_z  : one lane per (y,x) column, 64 columns per tile; z-forward,
//                 z-backward (with its interpolation step) and the clamp.
// Loop ranges reproduce the reference's exactly, including the partial
// coverage of rows/columns/layers outside the in/out intersection.
#include "gg_internal.h"

namespace gg {
namespace {
constexpr int kGridPerCU = 256;  // workgroups per CU of a redistancing launch

template <class T>
__device__ inline T absv(T v) { return v < T(0) ? -v : v; }

// One sweep step with interpolation on a sign change of the input values a
// (at the current point) and b (at the next point): may update cur, returns
// the candidate for the next point.
template <class T>
__device__ inline T step_interp(T a, T b, T& cur, T h) {
  if (a * b <= T(0)) {
    const T slope = (b - a) / h;
    const T d = -a / slope;
    if (absv(cur) > absv(d)) cur = -d * (T)sgnv(slope);
  }
  return cur + ((T)sgnv(b) * h);
}

template <class T, int NMAX>
__global__ __launch_bounds__(128) void k_redist_xy(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                   int32_t* err) {
  constexpr int S = NMAX + 1;
  __shared__ T O[NMAX * S];
  __shared__ T I[NMAX * S];
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* ain = (const T*)din.arena;
  T* aout = (T*)dout.arena;
  const T h = (T)sp.h;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, kl;
    plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* ob = dout.bounds + 6 * g;
    const int32_t* ib = din.bounds + 6 * g;
    const int ox0 = ob[0], oy0 = ob[1], oz0 = ob[2], ox1 = ob[3], oy1 = ob[4];
    const int no = ox1 - ox0;
    const int ni = ib[3] - ib[0];
    const int k = oz0 + kl;
    const int ixmin = max(ib[0], ox0), iymin = max(ib[1], oy0), izmin = max(ib[2], oz0);
    const int ixmax = min(ib[3], ox1), iymax = min(ib[4], oy1), izmax = min(ib[5], ob[5]);
    const bool empty = ixmin >= ixmax || iymin >= iymax || izmin >= izmax;
    if (empty && tid == 0 && kl == 0) atomicOr(err, kErrEmptyIO);
    for (int e = tid; e < no * no; e += nt) O[(e / no) * S + (e % no)] = T(-1);
    const bool do_xy = !empty && k >= izmin && k < izmax - 1;
    if (do_xy) {
      const int wI = ixmax - ixmin, hI = iymax - iymin;
      const T* src = ain + din.off[g] + (int64_t)(k - ib[2]) * ni * ni;
      stage_rows_to_lds(I + (iymin - oy0) * S + (ixmin - ox0), S, src + (int64_t)(iymin - ib[1]) * ni + (ixmin - ib[0]),
                        ni, hI, wI, tid, nt);
    }
    __syncthreads();
    if (do_xy) {
      // x-forward then x-backward, one lane per row i in [iymin, oy1).  The next
      // step's LDS values are read one step ahead (software pipelining); each
      // step still updates exactly the reference's values in its order.
      for (int r = tid; r < oy1 - iymin; r += nt) {
        const int i = iymin + r;
        T* orow = O + (i - oy0) * S;  // orow[j - ox0] = out(i, j, k)
        const T* irow = I + (i - oy0) * S;
        const bool inrow = i < iymax;
        if (ixmin < ox1 - 1) {
          int jj = ixmin - ox0;
          T cur = orow[jj], nxt = orow[jj + 1];
          T ia = irow[jj], ibv = irow[jj + 1];
          for (int j = ixmin; j < ox1 - 1; ++j, ++jj) {
            const bool more = j + 2 < ox1;
            const T nxt2 = more ? orow[jj + 2] : T(0);
            const T ib2 = more ? irow[jj + 2] : T(0);
            T cand;
            if (j < ixmax - 1 && inrow) cand = step_interp(ia, ibv, cur, h);
            else cand = cur + ((T)sgnv(nxt) * h);
            if (absv(cand) < absv(nxt)) nxt = cand;
            orow[jj] = cur;
            cur = nxt;
            nxt = nxt2;
            ia = ibv;
            ibv = ib2;
          }
          orow[ox1 - 1 - ox0] = cur;
        }
        if (ixmax - 1 > ox0) {
          int jj = ixmax - 1 - ox0;
          T cur = orow[jj], prv = orow[jj - 1], ip = irow[jj - 1];
          for (int j = ixmax - 1; j > ox0; --j, --jj) {
            const bool more = j - 2 >= ox0;
            const T prv2 = more ? orow[jj - 2] : T(0);
            const T ip2 = more ? irow[jj - 2] : T(0);
            T cand;
            if (j > ixmin && inrow) cand = cur + ((T)sgnv(ip) * h);
            else cand = cur + (T)sgnv(prv) * h;
            if (absv(cand) < absv(prv)) prv = cand;
            cur = prv;
            orow[jj - 1] = prv;
            prv = prv2;
            ip = ip2;
          }
        }
      }
      __syncthreads();
      // y-forward then y-backward, one lane per column j in [ixmin, ox1)
      for (int q = tid; q < ox1 - ixmin; q += nt) {
        const int j = ixmin + q;
        T* ocol = O + (j - ox0);  // ocol[(i - oy0) * S] = out(i, j, k)
        const T* icol = I + (j - ox0);
        const bool incol = j < ixmax;
        if (iymin < oy1 - 1) {
          int ii = iymin - oy0;
          T cur = ocol[ii * S], nxt = ocol[(ii + 1) * S];
          T ia = icol[ii * S], ibv = icol[(ii + 1) * S];
          for (int i = iymin; i < oy1 - 1; ++i, ++ii) {
            const bool more = i + 2 < oy1;
            const T nxt2 = more ? ocol[(ii + 2) * S] : T(0);
            const T ib2 = more ? icol[(ii + 2) * S] : T(0);
            T cand;
            if (incol && i < iymax - 1) cand = step_interp(ia, ibv, cur, h);
            else cand = cur + ((T)sgnv(nxt) * h);
            if (absv(cand) < absv(nxt)) nxt = cand;
            ocol[ii * S] = cur;
            cur = nxt;
            nxt = nxt2;
            ia = ibv;
            ibv = ib2;
          }
          ocol[(oy1 - 1 - oy0) * S] = cur;
        }
        if (iymax - 1 > oy0) {
          int ii = iymax - 1 - oy0;
          T cur = ocol[ii * S], prv = ocol[(ii - 1) * S], ip = icol[(ii - 1) * S];
          for (int i = iymax - 1; i > oy0; --i, --ii) {
            const bool more = i - 2 >= oy0;
            const T prv2 = more ? ocol[(ii - 2) * S] : T(0);
            const T ip2 = more ? icol[(ii - 2) * S] : T(0);
            T cand;
            if (incol && i > iymin) cand = cur + ((T)sgnv(ip) * h);
            else cand = cur + ((T)sgnv(prv) * h);
            if (absv(cand) < absv(prv)) prv = cand;
            cur = prv;
            ocol[(ii - 1) * S] = prv;
            prv = prv2;
            ip = ip2;
          }
        }
      }
    }
    __syncthreads();
    T* dst = aout + dout.off[g] + (int64_t)kl * no * no;
    for (int e = tid; e < no * no; e += nt) dst[e] = O[(e / no) * S + (e % no)];
    __syncthreads();
  }
}

// z-forward, z-backward and the clamp for 64 columns (one per lane).  Each lane
// stages its own out column and the in values it reads in LDS (only its own
// slots, so no barrier is needed): all global loads of the column are
// independent and coalesced across the wavefront, the sequential sweep then
// runs on LDS, and the clamped column is written back once.
template <class T, int NMAX>
__global__ __launch_bounds__(64) void k_redist_z(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout) {
  __shared__ T O[NMAX * 64];
  __shared__ T I[NMAX * 64];
  const int total = P.totals[kQCols * kClasses + cls];
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const T h = (T)sp.h;
  const T dl = (T)sp.delta;
  const int lane = threadIdx.x;
  T* Oc = O + lane;  // Oc[(k - oz0) * 64] = out(i, j, k)
  T* Ic = I + lane;  // Ic[(k - oz0) * 64] = in(i, j, k) where defined
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQCols, tile, g, blk);
    const int32_t* ob = dout.bounds + 6 * g;
    const int32_t* ib = din.bounds + 6 * g;
    const int ox0 = ob[0], oy0 = ob[1], oz0 = ob[2], oz1 = ob[5];
    const int no = ob[3] - ox0;
    const int ni = ib[3] - ib[0];
    const int col = blk * kColBlock + lane;
    if (col >= no * no) continue;
    const int i = oy0 + col / no, j = ox0 + col % no;
    const int ixmin = max(ib[0], ox0), iymin = max(ib[1], oy0), izmin = max(ib[2], oz0);
    const int ixmax = min(ib[3], ob[3]), iymax = min(ib[4], ob[4]), izmax = min(ib[5], oz1);
    const bool empty = ixmin >= ixmax || iymin >= iymax || izmin >= izmax;
    const int64_t ls = (int64_t)no * no;
    const int64_t ob_i = dout.off[g] + col;
    // batches of kStageBatch independent loads before their LDS writes
    for (int k0 = 0; k0 < no; k0 += kStageBatch) {
      T v[kStageBatch];
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u)
        if (k0 + u < no) v[u] = aout[ob_i + (k0 + u) * ls];
#pragma unroll
      for (int u = 0; u < kStageBatch; ++u)
        if (k0 + u < no) Oc[(k0 + u) * 64] = v[u];
    }
    if (!empty && i >= iymin && j >= ixmin) {
      const bool inside = i < iymax && j < ixmax;
      if (inside) {
        const int64_t lsi = (int64_t)ni * ni;
        const int64_t ib_i = din.off[g] + (int64_t)(i - ib[1]) * ni + (j - ib[0]);
        for (int k0 = izmin; k0 < izmax; k0 += kStageBatch) {
          T v[kStageBatch];
#pragma unroll
          for (int u = 0; u < kStageBatch; ++u)
            if (k0 + u < izmax) v[u] = ain[ib_i + (int64_t)(k0 + u - ib[2]) * lsi];
#pragma unroll
          for (int u = 0; u < kStageBatch; ++u)
            if (k0 + u < izmax) Ic[(k0 + u - oz0) * 64] = v[u];
        }
      }
#define OZ(k) Oc[((k) - oz0) * 64]
#define IZ(k) Ic[((k) - oz0) * 64]
      // software-pipelined: the next step's LDS values are read one step ahead
      if (izmin < oz1 - 1) {
        T cur = OZ(izmin), nxt = OZ(izmin + 1);
        T ia = IZ(izmin), ibv = IZ(izmin + 1);
        for (int k = izmin; k < oz1 - 1; ++k) {
          const bool more = k + 2 < oz1;
          const T nxt2 = more ? OZ(k + 2) : T(0);
          const T ib2 = more ? IZ(k + 2) : T(0);
          T cand;
          if (k < izmax - 1 && inside) cand = step_interp(ia, ibv, cur, h);
          else cand = cur + ((T)sgnv(nxt) * h);
          if (absv(cand) < absv(nxt)) nxt = cand;
          OZ(k) = cur;
          cur = nxt;
          nxt = nxt2;
          ia = ibv;
          ibv = ib2;
        }
        OZ(oz1 - 1) = cur;
      }
      if (izmax - 1 > oz0) {
        T cur = OZ(izmax - 1), prv = OZ(izmax - 2);
        T ic = IZ(izmax - 1), ip = IZ(izmax - 2);
        for (int k = izmax - 1; k > oz0; --k) {
          const bool more = k - 2 >= oz0;
          const T prv2 = more ? OZ(k - 2) : T(0);
          const T ip2 = more ? IZ(k - 2) : T(0);
          T cand;
          if (k > izmin && inside) cand = step_interp(ic, ip, cur, h);
          else cand = cur + ((T)sgnv(prv) * h);
          if (absv(cand) < absv(prv)) prv = cand;
          OZ(k) = cur;
          cur = prv;
          prv = prv2;
          ic = ip;
          ip = ip2;
        }
        OZ(oz0) = cur;
      }
#undef OZ
#undef IZ
    }
    // clampValues(-delta, delta) (src/dimensionalBufferReal.h:44-57)
    for (int k = 0; k < no; ++k) {
      T v = Oc[k * 64];
      if (v < -dl) v = -dl;
      else if (v > dl) v = dl;
      aout[ob_i + k * ls] = v;
    }
  }
}

// One forward step of a register-resident sweep at q -> q+1 (the loop bodies
// above): `act` = q inside the loop range, `ip` = interpolate (in values a, b
// at q, q+1).  FP64 keeps the reference's two divisions (bit-identical); FP32
// uses d = -a*h/(b - a) with the hardware reciprocal and no branch, which
// rounds differently but is the same formula.
__device__ inline void fwd_step(double& cur, double& nxt, double a, double b, bool act, bool ip, double h) {
  if (!act) return;
  double cand;
  if (ip) cand = step_interp(a, b, cur, h);
  else cand = cur + ((double)sgnv(nxt) * h);
  if (absv(cand) < absv(nxt)) nxt = cand;
}
__device__ inline float sgnh(float v, float h) { return v > 0.f ? h : (v < 0.f ? -h : 0.f); }
__device__ inline void fwd_step(float& cur, float& nxt, float a, float b, bool act, bool ip, float h) {
  const float dba = b - a;
  const float d = -a * h * __builtin_amdgcn_rcpf(dba);
  const bool moved = ip & (a * b <= 0.f) & (fabsf(cur) > fabsf(d));
  const float c1 = moved ? -d * (dba > 0.f ? 1.f : (dba < 0.f ? -1.f : 0.f)) : cur;
  const float cand = c1 + sgnh(ip ? b : nxt, h);
  const bool take = act & (fabsf(cand) < fabsf(nxt));
  cur = act ? c1 : cur;
  nxt = take ? cand : nxt;
}

// Register-resident forms of the two kernels (boxes up to 32 in both types,
// up to 48 in FP32): one wavefront per tile, a lane keeps its whole line (out
// and in values) in registers, the sweeps are unrolled over the line with the
// reference's loop ranges as predicates, and LDS is used once, to turn rows
// into columns.  The update order along a line is the same as above, so FP64
// results are identical; the FP32 steps are written without branches (vector
// selects), which a sequential recurrence needs to avoid an exec-mask round
// trip per step.

// x/y sweeps of one output layer.  LDS B first holds the in/out intersection of
// the input layer (read as rows), then the swept out rows (read as columns).
// Boxes up to 32 put two layers (two tiles) in one wavefront, 32 lanes each,
// so no lane idles for want of a row; what was uniform for a tile is then a
// per-lane value of its half.
//
// The sweeps are unrolled over NS <= NMAX positions, NS the smallest of NMAX -
// 12, - 8, - 4, NMAX that holds the wavefront's largest box side: positions
// past the side are inactive in every lane (the loop ranges end at no - 1), so
// a tile chooses its length once, by a uniform branch, and each length stays
// straight-line code.
struct XYTile {
  const int32_t* ib;
  const void* src;
  void* dst;
  int ni, ox0, oy0, no, iymin, iymax, ixmin, ixmax;
  bool do_xy, tv;
};

template <class T, int NMAX, int NS, int LW>
__device__ __forceinline__ void xy_sweeps(T* B, const XYTile& t, int xlo, int xim, int ylo, int yim, int nol, int ln,
                                          T h) {
  constexpr int S = NMAX + 1;
  const T* src = (const T*)t.src;
  const int no = t.no, ni = t.ni;
  T o[NS], in[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    o[q] = T(-1);
    __asm__ volatile("" : "+v"(o[q]));  // opaque: no control flow specialised on the initial value
  }
  // (the barriers sit outside the per-tile conditions: the two tiles of a
  // wavefront may differ in them)
  if (t.do_xy)
    stage_rows_to_lds(B + ylo * S + xlo, S, src + (int64_t)(t.iymin - t.ib[1]) * ni + (t.ixmin - t.ib[0]), ni,
                      t.iymax - t.iymin, t.ixmax - t.ixmin, ln, LW);
  __syncthreads();
  // x-forward then x-backward along row oy0 + ln, for rows in [iymin, oy1)
  const int r = ln;
  if (t.do_xy && r >= ylo && r < no) {
    const bool inrow = r <= yim;
#pragma unroll
    for (int q = 0; q < NS; ++q) in[q] = B[r * S + q];
#pragma unroll
    for (int q = 0; q < NS - 1; ++q)
      fwd_step(o[q], o[q + 1], in[q], in[q + 1], (q >= xlo) & (q < nol), (q < xim) & inrow, h);
#pragma unroll
    for (int q = NS - 1; q >= 1; --q) {
      const T cand = o[q] + ((T)sgnv(((q > xlo) & inrow) ? in[q - 1] : o[q - 1]) * h);
      o[q - 1] = ((q <= xim) & (absv(cand) < absv(o[q - 1]))) ? cand : o[q - 1];
    }
  }
  __syncthreads();
  if (t.do_xy && r < no) {
#pragma unroll
    for (int q = 0; q < NS; ++q) B[r * S + q] = o[q];
  }
  __syncthreads();
  if (t.do_xy) {
    // y-forward then y-backward along column ox0 + ln, for columns in [ixmin, ox1)
    const int c = ln;
    if (c < no) {
#pragma unroll
      for (int q = 0; q < NS; ++q) o[q] = B[q * S + c];
    }
    if (c >= xlo && c < no) {
      const bool incol = c <= xim;
      if (incol) {
        const T* scol = src + (t.ox0 + c - t.ib[0]) + (int64_t)(t.oy0 - t.ib[1]) * ni;  // scol[q * ni] = in(oy0 + q, j)
#pragma unroll
        for (int q = 0; q < NS; ++q)
          if (q >= ylo && q <= yim) in[q] = scol[(int64_t)q * ni];
      }
#pragma unroll
      for (int q = 0; q < NS - 1; ++q)
        fwd_step(o[q], o[q + 1], in[q], in[q + 1], (q >= ylo) & (q < nol), incol & (q < yim), h);
#pragma unroll
      for (int q = NS - 1; q >= 1; --q) {
        const T cand = o[q] + ((T)sgnv((incol & (q > ylo)) ? in[q - 1] : o[q - 1]) * h);
        o[q - 1] = ((q <= yim) & (absv(cand) < absv(o[q - 1]))) ? cand : o[q - 1];
      }
    }
  }
  // lane c writes column ox0 + c of the layer (coalesced across lanes)
  T* dst = (T*)t.dst;
  if (t.tv && ln < no) {
#pragma unroll
    for (int q = 0; q < NS; ++q)
      if (q < no) dst[q * no + ln] = o[q];
  }
  __syncthreads();
}

template <class T, int NMAX>
__global__ __launch_bounds__(64) void k_redist_xy_reg(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout,
                                                      int32_t* err) {
  constexpr int S = NMAX + 1;
  constexpr int PACK = NMAX <= 32 ? 2 : 1;  // tiles per wavefront
  constexpr int LW = 64 / PACK;             // lanes per tile
  __shared__ T Bs[PACK * NMAX * S];
  const int total = P.totals[kQLayers * kClasses + cls];
  const T* ain = (const T*)din.arena;
  T* aout = (T*)dout.arena;
  const T h = (T)sp.h;
  const int lane = threadIdx.x;
  const int sub = lane / LW, ln = lane % LW;  // tile of the wavefront, lane in the tile
  T* B = Bs + sub * NMAX * S;
  for (int tp = blockIdx.x; tp * PACK < total; tp += gridDim.x) {
    const int tile = tp * PACK + sub;
    const bool tv = tile < total;
    int g = 0, kl = 0;
    if (tv) plan_tile(P, cls, kQLayers, tile, g, kl);
    const int32_t* ob = dout.bounds + 6 * g;
    const int32_t* ib = din.bounds + 6 * g;
    const int ox0 = ob[0], oy0 = ob[1], oz0 = ob[2], ox1 = ob[3], oy1 = ob[4];
    const int no = ox1 - ox0;
    const int ni = ib[3] - ib[0];
    const int k = oz0 + kl;
    const int ixmin = max(ib[0], ox0), iymin = max(ib[1], oy0), izmin = max(ib[2], oz0);
    const int ixmax = min(ib[3], ox1), iymax = min(ib[4], oy1), izmax = min(ib[5], ob[5]);
    const bool empty = ixmin >= ixmax || iymin >= iymax || izmin >= izmax;
    if (tv && empty && ln == 0 && kl == 0) atomicOr(err, kErrEmptyIO);
    const bool do_xy = tv && !empty && k >= izmin && k < izmax - 1;
    const XYTile t{ib, ain + din.off[g] + (int64_t)(k - ib[2]) * ni * ni, aout + dout.off[g] + (int64_t)kl * no * no,
                   ni, ox0, oy0, no, iymin, iymax, ixmin, ixmax, do_xy, tv};
    // local (box) coordinates of the loop ranges
    // loop-range bounds as per-lane values: the predicates below then compare in
    // vector registers instead of building a scalar lane mask for every step
    int xlo = ixmin - ox0, xim = ixmax - 1 - ox0, ylo = iymin - oy0, yim = iymax - 1 - oy0, nol = no - 1;
    __asm__ volatile("" : "+v"(xlo), "+v"(xim), "+v"(ylo), "+v"(yim), "+v"(nol));
    // the wavefront's largest box side picks the sweep length
    int nw = tv ? no : 0;
    if constexpr (PACK == 2) nw = max(nw, __shfl_xor(nw, 32));
    nw = __builtin_amdgcn_readfirstlane(nw);
    if (nw <= NMAX - 12) xy_sweeps<T, NMAX, NMAX - 12, LW>(B, t, xlo, xim, ylo, yim, nol, ln, h);
    else if (nw <= NMAX - 8) xy_sweeps<T, NMAX, NMAX - 8, LW>(B, t, xlo, xim, ylo, yim, nol, ln, h);
    else if (nw <= NMAX - 4) xy_sweeps<T, NMAX, NMAX - 4, LW>(B, t, xlo, xim, ylo, yim, nol, ln, h);
    else xy_sweeps<T, NMAX, NMAX, LW>(B, t, xlo, xim, ylo, yim, nol, ln, h);
  }
}

// z-forward, z-backward and the clamp for 64 columns, one per lane, the column
// in registers.
// FP32 n <= 48: three waves per SIMD asked of the compiler (it fits the
// columns in 168 VGPRs without spilling; 20 % faster than the two it reaches
// unasked).
template <class T, int NMAX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && NMAX == 48 ? 3 : 1)))
void k_redist_z_reg(SimParams sp, PlanDev P, int cls, DescDev din, DescDev dout) {
  const int total = P.totals[kQCols * kClasses + cls];
  const T* __restrict__ ain = (const T*)din.arena;
  T* __restrict__ aout = (T*)dout.arena;
  const T h = (T)sp.h;
  const T dl = (T)sp.delta;
  const int lane = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQCols, tile, g, blk);
    const int32_t* ob = dout.bounds + 6 * g;
    const int32_t* ib = din.bounds + 6 * g;
    const int ox0 = ob[0], oy0 = ob[1], oz0 = ob[2], oz1 = ob[5];
    const int no = ob[3] - ox0;
    const int ni = ib[3] - ib[0];
    const int col = blk * kColBlock + lane;
    if (col >= no * no) continue;
    const int i = oy0 + col / no, j = ox0 + col % no;
    const int ixmin = max(ib[0], ox0), iymin = max(ib[1], oy0), izmin = max(ib[2], oz0);
    const int ixmax = min(ib[3], ob[3]), iymax = min(ib[4], ob[4]), izmax = min(ib[5], oz1);
    const bool empty = ixmin >= ixmax || iymin >= iymax || izmin >= izmax;
    const int64_t ls = (int64_t)no * no;
    T* oc = aout + dout.off[g] + col;  // oc[q * ls] = out(i, j, oz0 + q)
    T o[NMAX], in[NMAX];
#pragma unroll
    for (int q = 0; q < NMAX; ++q)
      if (q < no) o[q] = oc[q * ls];
    if (!empty && i >= iymin && j >= ixmin) {
      const bool inside = i < iymax && j < ixmax;
      int zlo = izmin - oz0, zim = izmax - 1 - oz0, nol = no - 1;
      __asm__ volatile("" : "+v"(zlo), "+v"(zim), "+v"(nol));  // per-lane predicates (see above)
      if (inside) {
        const int64_t lsi = (int64_t)ni * ni;
        const T* ic = ain + din.off[g] + (int64_t)(i - ib[1]) * ni + (j - ib[0]) + (int64_t)(oz0 - ib[2]) * lsi;
#pragma unroll
        for (int q = 0; q < NMAX; ++q)
          if (q >= zlo && q <= zim) in[q] = ic[q * lsi];
      }
#pragma unroll
      for (int q = 0; q < NMAX - 1; ++q)
        fwd_step(o[q], o[q + 1], in[q], in[q + 1], (q >= zlo) & (q < nol), (q < zim) & inside, h);
      // z-backward interpolates too: the same step towards q - 1
#pragma unroll
      for (int q = NMAX - 1; q >= 1; --q)
        fwd_step(o[q], o[q - 1], in[q], in[q - 1], q <= zim, (q > zlo) & inside, h);
    }
    // clampValues(-delta, delta) (src/dimensionalBufferReal.h:44-57)
#pragma unroll
    for (int q = 0; q < NMAX; ++q) {
      if (q < no) {
        T v = o[q];
        if (v < -dl) v = -dl;
        else if (v > dl) v = dl;
        oc[q * ls] = v;
      }
    }
  }
}

template <class T, int NMAX>
void xy_class(const Launch& L, const SimParams& sp, const PlanDev& P, int cls, const DescDev& din,
              const DescDev& dout, int32_t* err) {
  hipLaunchKernelGGL((k_redist_xy<T, NMAX>), dim3(tile_grid(L, kGridPerCU, cls)), dim3(128), 0, L.stream, sp, P, cls, din, dout, err);
}

// part 0: x/y sweeps per layer (one launch per size class); part 1: z sweeps + clamp
template <class T>
void redist_all(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                const DescDev& dout, int32_t* err, int c0, int c1) {
  // register-resident lines for boxes up to 32 (both types) and up to 48 (FP32)
  constexpr bool reg48 = sizeof(T) == 4;
  auto in = [&](int c) { return c0 <= c && c < c1; };
  if (part == 0) {
    if (in(0))
      hipLaunchKernelGGL((k_redist_xy_reg<T, 32>), dim3(tile_grid(L, kGridPerCU, 0)), dim3(64), 0, L.stream, sp, P, 0,
                         din, dout, err);
    if (in(1)) {
      if constexpr (reg48)
        hipLaunchKernelGGL((k_redist_xy_reg<T, 48>), dim3(tile_grid(L, kGridPerCU, 1)), dim3(64), 0, L.stream, sp, P,
                           1, din, dout, err);
      else
        xy_class<T, 48>(L, sp, P, 1, din, dout, err);
    }
    if (in(2)) xy_class<T, 64>(L, sp, P, 2, din, dout, err);
    if (in(3)) xy_class<T, 96>(L, sp, P, 3, din, dout, err);
  } else {
    if (in(0))
      hipLaunchKernelGGL((k_redist_z_reg<T, 32>), dim3(tile_grid(L, kGridPerCU, 0)), dim3(64), 0, L.stream, sp, P, 0,
                         din, dout);
    if (in(1)) {
      if constexpr (reg48)
        hipLaunchKernelGGL((k_redist_z_reg<T, 48>), dim3(tile_grid(L, kGridPerCU, 1)), dim3(64), 0, L.stream, sp, P, 1,
                           din, dout);
      else
        hipLaunchKernelGGL((k_redist_z<T, 48>), dim3(tile_grid(L, kGridPerCU, 1)), dim3(64), 0, L.stream, sp, P, 1,
                           din, dout);
    }
    if (in(2))
      hipLaunchKernelGGL((k_redist_z<T, 64>), dim3(tile_grid(L, kGridPerCU, 2)), dim3(64), 0, L.stream, sp, P, 2, din,
                         dout);
    if (in(3))
      hipLaunchKernelGGL((k_redist_z<T, 96>), dim3(tile_grid(L, kGridPerCU, 3)), dim3(64), 0, L.stream, sp, P, 3, din,
                         dout);
  }
}

}  // namespace

void launch_redist(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                   const DescDev& dout, int32_t* err, bool fp32, int c0, int c1) {
  if (fp32) redist_all<float>(L, part, sp, P, din, dout, err, c0, c1);
  else redist_all<double>(L, part, sp, P, din, dout, err, c0, c1);
}

}  // namespace gg
