// Heat-kernel convolution of every grain box (LSbox::executeConvolution,
// src/box.cpp:414-539, Gaussian branch of convolutionGeneratorMKL :751-767).
//
// FP64 plain form: mixed-radix FFT (radices 4, 2, 3, 5 specialised, any other
// prime factor as a direct sum) on the n^3 box, half spectrum along x (two
// real rows per complex FFT), spectral factor G, inverse FFT.  The library is
// built with -ffp-contract=off and every operation follows this fixed
// definition, so each value is the same IEEE double the FP64 CPU
// restatement produces:
//   Stockham pass of radix p (T = n/p, Ns = product of the earlier radices;
//   radices 4 while 4 divides, then 2, then odd primes ascending): for
//   i < T, j = i mod Ns: v_r = x[i + r*T]; if j > 0, v_r = twiddle(v_r,
//   r*j*n/(Ns*p)) for r >= 1, forward (a*c + b*s, b*c - a*s), inverse
//   (a*c - b*s, a*s + b*c), c/s = cos/sin(2 pi t/n) host tables;
//   y = butterfly_p(v); x'[(i-j)*p + j + k*Ns] = y_k.  Butterflies as in
//   butterfly() below; other (odd prime) radices as conjugate-pair sums
//   (fft_pass_any).
//   x forward: rows 2q, 2q+1 as a + ib; S[2q][k] = (Z_k + conj Z_(n-k))/2,
//   S[2q+1][k] = (Z_k - conj Z_(n-k))/2i, written as in k_fft_xy_fwd.
//   G: re*G, im*G.  x inverse: Z_k = A_k + i B_k with the Hermitian
//   extension for k > n/2; row 2q = Re IFFT(Z), row 2q+1 = Im.
// Three launches per size class: (x,y) forward per z-layer with the layer in
// LDS, z forward + G + z inverse per block of 32 spectral columns, (y,x)
// inverse per z-layer.
//
// FP32 batched forms: the Gaussian factor is separable,
// G(kx,ky,kz) = f(kx) f(ky) f(kz), so IDFT . diag(G) . DFT on the box equals
// three 1-D circular convolutions with the real kernel
// c[m] = (1/n) sum_k f(min(k,n-k)) cos(2 pi k m / n) (host long double, stored
// FP32).  Two launches per size class: x then y per z-layer in LDS (vector
// FMAs), then z per block of 64 columns (matrix cores, FP32); or, as the
// split-BF16 form, all three passes as GEMMs on the matrix cores with 3-term
// BF16 products accumulated in FP32.
#include "gg_internal.h"

namespace gg {
namespace {
constexpr int kGridPerCU = 128;  // workgroups per CU of a convolution launch

constexpr int kT = 256;

// ---- FP64 mixed-radix FFT, staged in LDS ------------------------------------
// A batch of nl lines of length n sits in LDS (element e of line l at
// buf[l*ls + e*es]).  Each Stockham pass (radix p, T = n/p butterflies per
// line) reads every butterfly's p inputs and keeps its p outputs in registers,
// and after a barrier writes them to their sorted places, so one buffer
// serves all passes.  Consecutive lanes take the same butterfly of
// consecutive lines: columns (ls = 1) are then read and written in contiguous
// runs, and rows are laid out at an odd stride (in complex values), so the
// lanes of a 16-byte access fall on distinct bank groups.
constexpr double kC1 = 0.309016994374947424102293417182819059;
constexpr double kC2 = -0.809016994374947424102293417182819059;
constexpr double kS1 = 0.951056516295153572116439333379382143;
constexpr double kS2 = 0.587785252292473129168705954639072769;
constexpr double kS3 = 0.866025403784438646763723170752936183;
constexpr int kPlanLen = 8;  // radices per length in the plan table (0-terminated)

__host__ __device__ constexpr int cdiv(int a, int b) { return (a + b - 1) / b; }

// a / d for 0 <= a < 2^20, 1 <= d <= 2^10, inv = 1.0f / d
__device__ inline int sdiv(int a, int d, float inv) {
  int q = (int)((float)a * inv);
  const int r = a - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

template <bool INV>
__device__ inline double2 twid(double2 v, double c, double s) {
  if (INV) return make_double2(v.x * c - v.y * s, v.x * s + v.y * c);
  return make_double2(v.x * c + v.y * s, v.y * c - v.x * s);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

template <int P, bool INV>
__device__ inline void butterfly(const double2 (&v)[P], double2 (&y)[P]) {
  if constexpr (P == 2) {
    y[0] = cadd(v[0], v[1]);
    y[1] = csub(v[0], v[1]);
  } else if constexpr (P == 4) {
    const double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]), a2 = cadd(v[1], v[3]), a3 = csub(v[1], v[3]);
    y[0] = cadd(a0, a2);
    y[2] = csub(a0, a2);
    const double2 m1 = make_double2(a1.x + a3.y, a1.y - a3.x), p1 = make_double2(a1.x - a3.y, a1.y + a3.x);
    y[1] = INV ? p1 : m1;
    y[3] = INV ? m1 : p1;
  } else if constexpr (P == 3) {
    const double2 t1 = cadd(v[1], v[2]);
    const double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
    const double2 t3 = make_double2(kS3 * (v[1].x - v[2].x), kS3 * (v[1].y - v[2].y));
    y[0] = cadd(v[0], t1);
    const double2 m1 = make_double2(t2.x + t3.y, t2.y - t3.x), p1 = make_double2(t2.x - t3.y, t2.y + t3.x);
    y[1] = INV ? p1 : m1;
    y[2] = INV ? m1 : p1;
  } else {
    static_assert(P == 5, "specialised radices: 2, 3, 4, 5");
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    y[0] = make_double2((v[0].x + t1.x) + t2.x, (v[0].y + t1.y) + t2.y);
    const double2 b1 = make_double2((v[0].x + kC1 * t1.x) + kC2 * t2.x, (v[0].y + kC1 * t1.y) + kC2 * t2.y);
    const double2 b2 = make_double2((v[0].x + kC2 * t1.x) + kC1 * t2.x, (v[0].y + kC2 * t1.y) + kC1 * t2.y);
    const double2 d1 = make_double2(kS1 * t3.x + kS2 * t4.x, kS1 * t3.y + kS2 * t4.y);
    const double2 d2 = make_double2(kS2 * t3.x - kS1 * t4.x, kS2 * t3.y - kS1 * t4.y);
    const double2 m1 = make_double2(b1.x + d1.y, b1.y - d1.x), p1 = make_double2(b1.x - d1.y, b1.y + d1.x);
    const double2 m2 = make_double2(b2.x + d2.y, b2.y - d2.x), p2 = make_double2(b2.x - d2.y, b2.y + d2.x);
    y[1] = INV ? p1 : m1;
    y[4] = INV ? m1 : p1;
    y[2] = INV ? p2 : m2;
    y[3] = INV ? m2 : p2;
  }
}

// One pass of a specialised radix P over nl lines (MAXOUT >= nl*n).
template <int P, bool INV, int MAXOUT>
__device__ inline void fft_pass(double2* buf, int n, int nl, int es, int ls, int Ns, const double* tc,
                                const double* ts, int tid) {
  constexpr int U = cdiv(cdiv(MAXOUT, P), kT);
  const int T = n / P, units = nl * T, step = n / (Ns * P);
  const float inv_nl = 1.0f / (float)nl, inv_Ns = 1.0f / (float)Ns;
  double2 y[U][P];
  int ob[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int w = tid + u * kT;
    ob[u] = -1;
    if (w < units) {
      const int i = sdiv(w, nl, inv_nl), line = w - i * nl;
      const int j = i - sdiv(i, Ns, inv_Ns) * Ns;
      const double2* src = buf + line * ls;
      double2 v[P];
#pragma unroll
      for (int r = 0; r < P; ++r) v[r] = src[(i + r * T) * es];
      if (j > 0) {
#pragma unroll
        for (int r = 1; r < P; ++r) {
          const int t = r * j * step;
          v[r] = twid<INV>(v[r], tc[t], ts[t]);
        }
      }
      butterfly<P, INV>(v, y[u]);
      ob[u] = line * ls + ((i - j) * P + j) * es;
    }
  }
  __syncthreads();
  const int os = Ns * es;
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (ob[u] >= 0) {
#pragma unroll
      for (int k = 0; k < P; ++k) buf[ob[u] + k * os] = y[u][k];
    }
  __syncthreads();
}

// One pass of an odd prime radix p > 5, h = (p-1)/2: work item k = 0..h of a
// butterfly makes y_0 (k = 0) or the conjugate pair y_k, y_(p-k) from the sums
// over r = 1..h of a_r = v_r + v_(p-r) (times cos) and b_r = v_r - v_(p-r)
// (times sin); consecutive lanes take the same item of consecutive lines.
template <bool INV, int MAXOUT>
__device__ inline void fft_pass_any(double2* buf, int n, int nl, int es, int ls, int Ns, int p, const double* tc,
                                    const double* ts, int tid) {
  constexpr int U = cdiv(cdiv(MAXOUT * 4, 7), kT);  // (h+1)/p <= 4/7 for p >= 7
  const int T = n / p, step = n / (Ns * p), h = (p - 1) / 2, K = h + 1;
  const int items = nl * T * K;
  const float inv_K = 1.0f / (float)K, inv_nl = 1.0f / (float)nl, inv_Ns = 1.0f / (float)Ns;
  double2 ya[U], yb[U];
  int oa[U], ob[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int w = tid + u * kT;
    oa[u] = -1;
    ob[u] = -1;
    if (w < items) {
      const int rest = sdiv(w, nl, inv_nl), line = w - rest * nl;
      const int i = sdiv(rest, K, inv_K), k = rest - i * K;
      const int j = i - sdiv(i, Ns, inv_Ns) * Ns;
      const double2* src = buf + line * ls;
      const double2 v0 = src[i * es];
      auto input = [&](int r) {
        double2 v = src[(i + r * T) * es];
        if (j > 0) {
          const int t = r * j * step;
          v = twid<INV>(v, tc[t], ts[t]);
        }
        return v;
      };
      const int base = line * ls + (i - j) * p * es + j * es;
      if (k == 0) {
        double2 y0 = v0;
        for (int r = 1; r <= h; ++r) {
          const double2 a = input(r), b = input(p - r);
          y0.x = y0.x + (a.x + b.x);
          y0.y = y0.y + (a.y + b.y);
        }
        ya[u] = y0;
        oa[u] = base;
      } else {
        double Pr = 0.0, Pi = 0.0, Qr = 0.0, Qi = 0.0;
        int rk = k;  // r*k mod p
        for (int r = 1; r <= h; ++r) {
          const double2 vr = input(r), vq = input(p - r);
          const double2 a = cadd(vr, vq), b = csub(vr, vq);
          const int t = rk * T;
          const double c = tc[t], s = ts[t];
          Pr = Pr + a.x * c;
          Pi = Pi + a.y * c;
          Qr = Qr + b.x * s;
          Qi = Qi + b.y * s;
          rk += k;
          if (rk >= p) rk -= p;
        }
        const double mr = v0.x + Pr, mi = v0.y + Pi;
        const double2 lo = make_double2(mr + Qi, mi - Qr), hi = make_double2(mr - Qi, mi + Qr);
        ya[u] = INV ? hi : lo;
        yb[u] = INV ? lo : hi;
        oa[u] = base + k * Ns * es;
        ob[u] = base + (p - k) * Ns * es;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < U; ++u) {
    if (oa[u] >= 0) buf[oa[u]] = ya[u];
    if (ob[u] >= 0) buf[ob[u]] = yb[u];
  }
  __syncthreads();
}

// The whole FFT of nl lines: the passes of n's plan in order.
template <bool INV, int MAXOUT>
__device__ inline void fft_lines(double2* buf, int n, int nl, int es, int ls, const int (&radix)[kPlanLen],
                                 const double* tc, const double* ts, int tid) {
  int Ns = 1;
  for (int s = 0; s < kPlanLen; ++s) {
    const int p = radix[s];
    if (p == 0) break;
    if (p == 4) fft_pass<4, INV, MAXOUT>(buf, n, nl, es, ls, Ns, tc, ts, tid);
    else if (p == 2) fft_pass<2, INV, MAXOUT>(buf, n, nl, es, ls, Ns, tc, ts, tid);
    else if (p == 3) fft_pass<3, INV, MAXOUT>(buf, n, nl, es, ls, Ns, tc, ts, tid);
    else if (p == 5) fft_pass<5, INV, MAXOUT>(buf, n, nl, es, ls, Ns, tc, ts, tid);
    else fft_pass_any<INV, MAXOUT>(buf, n, nl, es, ls, Ns, p, tc, ts, tid);
    Ns *= p;
  }
}

__device__ inline void load_fft_tables(const TablesDev& tb, int n, double* tc, double* ts, int (&radix)[kPlanLen],
                                       int tid) {
  const int to = tb.tw_off[n];
  for (int i = tid; i < n; i += kT) {
    tc[i] = tb.twc[to + i];
    ts[i] = tb.tws[to + i];
  }
#pragma unroll
  for (int s = 0; s < kPlanLen; ++s) radix[s] = tb.fplan[n * kPlanLen + s];
}

// (x, y) forward per z-layer: two real rows per complex x-FFT, the half
// spectrum separated from it, then the y-FFT of the n/2+1 spectral columns.
template <int NMAX>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(NMAX <= 32 ? 5 : NMAX <= 48 ? 4 : 1))) void k_fft_xy_fwd(PlanDev P, int cls, DescDev din, TablesDev tb,
                                                   const int64_t* __restrict__ spec_off, double2* spec) {
  constexpr int NR = (NMAX + 1) / 2, N2 = NMAX / 2 + 1;
  constexpr int R2 = cdiv(NMAX * N2, kT);
  static_assert(NR * (NMAX | 1) <= NMAX * N2, "padded rows fit the layer buffer");
  __shared__ double2 L[NMAX * N2];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQLayers * kClasses + cls];
  const double* arena = (const double*)din.arena;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(din.bounds, g), n2 = n / 2 + 1, nr = (n + 1) / 2, nc = n * n2;
    int radix[kPlanLen];
    load_fft_tables(tb, n, tc, ts, radix, tid);
    const double* src = arena + din.off[g] + (int64_t)z * n * n;
    const int rs = n | 1;  // row stride of the x-FFT lines (odd: lanes on consecutive rows hit distinct banks)
    for (int o = tid; o < nr * n; o += kT) {
      const int q = o / n, x = o - q * n;
      const double a = src[2 * q * n + x];
      const double b = 2 * q + 1 < n ? src[(2 * q + 1) * n + x] : 0.0;
      L[q * rs + x] = make_double2(a, b);
    }
    __syncthreads();
    fft_lines<false, NR * NMAX>(L, n, nr, 1, rs, radix, tc, ts, tid);
    // separate the two rows' half spectra: S[y][k], y = 2q (+1)
    double2 sv[R2];
#pragma unroll
    for (int u = 0; u < R2; ++u) {
      const int o = tid + u * kT;
      if (o < nc) {
        const int y = o / n2, k = o - y * n2, q = y >> 1, m = k == 0 ? 0 : n - k;
        const double2 zk = L[q * rs + k], zm = L[q * rs + m];
        sv[u] = (y & 1) ? make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x))
                        : make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < R2; ++u) {
      const int o = tid + u * kT;
      if (o < nc) L[o] = sv[u];
    }
    __syncthreads();
    fft_lines<false, N2 * NMAX>(L, n, n2, n2, 1, radix, tc, ts, tid);
    double2* dst = spec + spec_off[g] + (int64_t)z * nc;
    for (int o = tid; o < nc; o += kT) dst[o] = L[o];
    __syncthreads();
  }
}

// z forward, spectral factor, z inverse per block of kSpecBlock spectral columns.
template <int NMAX>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(NMAX <= 32 ? 5 : NMAX <= 48 ? 4 : 1))) void k_fft_z(PlanDev P, int cls, DescDev dref, TablesDev tb,
                                              const int64_t* __restrict__ spec_off, double2* spec) {
  constexpr int CB = kSpecBlock;
  __shared__ double2 B[NMAX * CB];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQSpecCols * kClasses + cls];
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQSpecCols, tile, g, blk);
    const int n = box_n(dref.bounds, g), n2 = n / 2 + 1, nc = n * n2;
    const int c0 = blk * CB, width = min(CB, nc - c0);
    const double* G = tb.gtab + tb.g_off[n];
    int radix[kPlanLen];
    load_fft_tables(tb, n, tc, ts, radix, tid);
    double2* base = spec + spec_off[g];
    stage_rows_to_lds(B, CB, base + c0, nc, n, width, tid, kT);
    __syncthreads();
    fft_lines<false, CB * NMAX>(B, n, width, CB, 1, radix, tc, ts, tid);
    for (int o = tid; o < n * CB; o += kT) {
      const int kz = o / CB, cb = o - kz * CB, c = c0 + cb;
      if (cb < width) {
        const int ky = c / n2, kx = c - ky * n2;
        const int kz2 = min(kz, n - kz), ky2 = min(ky, n - ky), kx2 = min(kx, n - kx);
        const double gf = G[kz2 * kz2 + ky2 * ky2 + kx2 * kx2];
        const double2 v = B[o];
        B[o] = make_double2(v.x * gf, v.y * gf);
      }
    }
    __syncthreads();
    fft_lines<true, CB * NMAX>(B, n, width, CB, 1, radix, tc, ts, tid);
    for (int o = tid; o < n * width; o += kT) {
      const int z = sdiv(o, width, 1.0f / (float)width), cb = o - z * width;
      base[(int64_t)z * nc + c0 + cb] = B[z * CB + cb];
    }
    __syncthreads();
  }
}

// (y, x) inverse per z-layer: y-IFFT of the spectral columns, two rows' half
// spectra packed into one full spectrum, x-IFFT, real and imaginary parts out.
template <int NMAX>
__global__ __launch_bounds__(kT) __attribute__((amdgpu_waves_per_eu(NMAX <= 32 ? 5 : NMAX <= 48 ? 4 : 1))) void k_fft_xy_inv(PlanDev P, int cls, DescDev dout, TablesDev tb,
                                                   const int64_t* __restrict__ spec_off, const double2* spec) {
  constexpr int NR = (NMAX + 1) / 2, N2 = NMAX / 2 + 1;
  constexpr int R1 = cdiv(NR * NMAX, kT);
  __shared__ double2 L[NMAX * N2];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQLayers * kClasses + cls];
  double* arena = (double*)dout.arena;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(dout.bounds, g), n2 = n / 2 + 1, nr = (n + 1) / 2, nc = n * n2;
    int radix[kPlanLen];
    load_fft_tables(tb, n, tc, ts, radix, tid);
    const double2* src = spec + spec_off[g] + (int64_t)z * nc;
    stage_to_lds(L, src, nc, tid, kT);
    __syncthreads();
    fft_lines<true, N2 * NMAX>(L, n, n2, n2, 1, radix, tc, ts, tid);
    double2 zv[R1];
#pragma unroll
    for (int u = 0; u < R1; ++u) {
      const int o = tid + u * kT;
      if (o < nr * n) {
        const int q = o / n, k = o - q * n;
        const bool lo = k < n2, two = 2 * q + 1 < n;
        const int m = lo ? k : n - k;
        const double2 A = L[2 * q * n2 + m];
        const double2 Bv = two ? L[(2 * q + 1) * n2 + m] : make_double2(0.0, 0.0);
        zv[u] = lo ? make_double2(A.x - Bv.y, A.y + Bv.x) : make_double2(A.x + Bv.y, Bv.x - A.y);
      }
    }
    __syncthreads();
    const int rs = n | 1;  // row stride of the x-FFT lines, as in the forward kernel
#pragma unroll
    for (int u = 0; u < R1; ++u) {
      const int o = tid + u * kT;
      if (o < nr * n) {
        const int q = o / n;
        L[o + q * (rs - n)] = zv[u];
      }
    }
    __syncthreads();
    fft_lines<true, NR * NMAX>(L, n, nr, 1, rs, radix, tc, ts, tid);
    double* dst = arena + dout.off[g] + (int64_t)z * n * n;
    for (int o = tid; o < nr * n; o += kT) {
      const int q = o / n, x = o - q * n;
      const double2 v = L[q * rs + x];
      dst[2 * q * n + x] = v.x;
      if (2 * q + 1 < n) dst[(2 * q + 1) * n + x] = v.y;
    }
    __syncthreads();
  }
}

// ---- FP32 separable circulant form ----------------------------------------
// Each lane produces 4 consecutive outputs of a line.  The kernel row is held
// twice over (crr[i] = c[i mod n], i < 2n+4) so (o - m) mod n needs no wrap, and
// the four coefficients slide through registers: per source value, one LDS
// read of the data, one of the kernel row, four FMAs.
constexpr int kRB = 4;

__device__ inline void load_circ_row(float* crr, const TablesDev& tb, int n, int tid, int nt) {
  const float* c = tb.circ + tb.tw_off[n];
  for (int i = tid; i < 2 * n + kRB; i += nt) crr[i] = c[i % n];
}

// out[o] = sum_m c[(o - m) mod n] * src[m*stride], o = o0 .. o0+3
__device__ inline void circ4(const float* crr, const float* src, int stride, int n, int o0, float acc[kRB]) {
  float w0 = crr[o0 + n], w1 = crr[o0 + 1 + n], w2 = crr[o0 + 2 + n], w3 = crr[o0 + 3 + n];
  acc[0] = acc[1] = acc[2] = acc[3] = 0.f;
  // unrolled by 4 so that the weight window rotates through register names
  // instead of three moves per tap
#pragma unroll 4
  for (int m = 0; m < n; ++m) {
    const float v = src[m * stride];
    acc[0] = __builtin_fmaf(w0, v, acc[0]);
    acc[1] = __builtin_fmaf(w1, v, acc[1]);
    acc[2] = __builtin_fmaf(w2, v, acc[2]);
    acc[3] = __builtin_fmaf(w3, v, acc[3]);
    w3 = w2;
    w2 = w1;
    w1 = w0;
    w0 = crr[o0 + n - m - 1];
  }
}

template <int NMAX>
__global__ __launch_bounds__(kT) void k_circ_xy(PlanDev P, int cls, DescDev din, DescDev dout, TablesDev tb) {
  __shared__ float A[NMAX * NMAX];
  __shared__ float B[NMAX * NMAX];
  __shared__ float crr[2 * NMAX + kRB];
  const int total = P.totals[kQLayers * kClasses + cls];
  const float* ain = (const float*)din.arena;
  float* aout = (float*)dout.arena;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(din.bounds, g), nn = n * n;
    const int G = (n + kRB - 1) / kRB;  // groups of 4 outputs per line
    load_circ_row(crr, tb, n, tid, kT);
    stage_to_lds(A, ain + din.off[g] + (int64_t)z * nn, nn, tid, kT);
    __syncthreads();
    // along x: item = (y, group)
    for (int it = tid; it < n * G; it += kT) {
      const int y = it / G, x0 = (it - y * G) * kRB;
      float acc[kRB];
      circ4(crr, A + y * n, 1, n, x0, acc);
#pragma unroll
      for (int u = 0; u < kRB; ++u)
        if (x0 + u < n) B[y * n + x0 + u] = acc[u];
    }
    __syncthreads();
    // along y: item = (group, x), consecutive lanes on consecutive x
    float* dst = aout + dout.off[g] + (int64_t)z * nn;
    for (int it = tid; it < n * G; it += kT) {
      const int gy = it / n, x = it - gy * n, y0 = gy * kRB;
      float acc[kRB];
      circ4(crr, B + x, n, n, y0, acc);
#pragma unroll
      for (int u = 0; u < kRB; ++u)
        if (y0 + u < n) dst[(y0 + u) * n + x] = acc[u];
    }
    __syncthreads();
  }
}

// Along z the convolution of a block of 64 columns is a product with the n x n
// circulant matrix C[z][k] = c[(z - k) mod n], a small GEMM run on the matrix
// cores (v_mfma_f32_16x16x4_f32, FP32 in and out).  n is padded to
// NP = 16*ceil(n/16); padded k contribute zero (both operands are masked, so
// stale LDS never meets a zero weight).  C is never stored: its entries are
// read from the kernel row held twice over (crr), at index (z - k) + n.
//
// Operand layout of the 16x16x4 instruction (64 lanes): lane l supplies
// A[l % 16][k0 + l / 16] and B[k0 + l / 16][l % 16]; it receives
// D[4 * (l / 16) + v][l % 16] in accumulator v = 0..3.
typedef float v4f __attribute__((ext_vector_type(4)));

template <int NMAX>
__global__ __launch_bounds__(kT) void k_circ_z(PlanDev P, int cls, DescDev dout, TablesDev tb) {
  constexpr int CB = kColBlock;
  constexpr int NPM = (NMAX + 15) / 16 * 16;
  constexpr int KSM = NPM / 4;  // k-steps of the largest box of the class
  constexpr int SZ = CB + 1;    // row stride of the staged column block
  __shared__ float Lz[NPM * SZ];
  __shared__ float crr[2 * NMAX + kRB];
  const int total = P.totals[kQCols * kClasses + cls];
  float* aout = (float*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQCols, tile, g, blk);
    const int n = box_n(dout.bounds, g), nn = n * n;
    const int c0 = blk * CB, width = min(CB, nn - c0);
    const int T16 = (n + 15) >> 4, KS = T16 * 4;
    load_circ_row(crr, tb, n, tid, kT);
    float* base = aout + dout.off[g];
    stage_rows_to_lds(Lz, SZ, base + c0, nn, n, width, tid, kT);
    __syncthreads();
    // out[z][p] = sum_k c[(z - k) mod n] Lz[k][p]; tiles of 16 z by 16 columns
    for (int t = wv; t < T16 * (CB / 16); t += kT / 64) {
      const int zt = t / (CB / 16), pt = t - zt * (CB / 16);
      const int za = zt * 16 + lr, pb = pt * 16 + lr;
      float av[KSM], bv[KSM];
#pragma unroll
      for (int kk = 0; kk < KSM; ++kk) {
        const int k = kk * 4 + lk;
        const bool ca = k < n && za < n;
        av[kk] = (kk < KS && ca) ? crr[ca ? za - k + n : 0] : 0.f;
        bv[kk] = (kk < KS && k < n) ? Lz[k * SZ + pb] : 0.f;
      }
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < KSM; ++kk)
        if (kk < KS) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kk], bv[kk], acc, 0, 0, 0);
      if (pb < width) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int zz = zt * 16 + 4 * lk + v;
          if (zz < n) base[(int64_t)zz * nn + c0 + pb] = acc[v];
        }
      }
    }
    __syncthreads();
  }
}

// ---- split-BF16 (3-term) circulant on the matrix cores ----------------------
// The same three circulant products as above, each as a GEMM on
// v_mfma_f32_16x16x32_bf16: every FP32 operand v is split into hi = bf16(v)
// (nearest even) and lo = bf16(v - hi), and a 32-deep k block adds
// a_lo b_hi + a_hi b_lo + a_hi b_hi (smallest terms first) to an FP32
// accumulator; the dropped a_lo b_lo term and the rounding of lo leave a
// relative error of about 2^-16 per product.  Operand layout: lane l supplies
// A[l % 16][k] and B[k][l % 16] for k = 32 kb + 8 (l / 16) + j, j = 0..7
// (the same k slots in both operands, so the product does not depend on how
// the hardware orders them), and receives D[4 (l / 16) + v][l % 16].
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ inline v4f mfma_split3(const float (&a)[8], const float (&b)[8], v4f acc) {
  v8bf ah, al, bh, bl;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 x = (__bf16)a[j], y = (__bf16)b[j];
    ah[j] = x;
    al[j] = (__bf16)(a[j] - (float)x);
    bh[j] = y;
    bl[j] = (__bf16)(b[j] - (float)y);
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
  return acc;
}

// x then y on a z-layer: D[x][y] = sum_k C[x][k] L[y][k], then
// D[y][x] = sum_k C[y][k] B[k][x]; layer rows at stride n + 1 in LDS.
template <int NMAX>
__global__ __launch_bounds__(kT) void k_circ_xy_bf16(PlanDev P, int cls, DescDev din, DescDev dout, TablesDev tb) {
  constexpr int RS = NMAX + 1;
  constexpr int KBM = (NMAX + 31) / 32;
  __shared__ float A[NMAX * RS];
  __shared__ float B[NMAX * RS];
  __shared__ float crr[2 * NMAX + kRB];
  const int total = P.totals[kQLayers * kClasses + cls];
  const float* ain = (const float*)din.arena;
  float* aout = (float*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lg = lane >> 4;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(din.bounds, g), nn = n * n, rs = n + 1;
    const int T16 = (n + 15) >> 4, KB = (n + 31) >> 5;
    load_circ_row(crr, tb, n, tid, kT);
    const float* src = ain + din.off[g] + (int64_t)z * nn;
    for (int i = tid; i < nn; i += kT) {
      const int y = i / n;
      A[y * rs + (i - y * n)] = src[i];
    }
    __syncthreads();
    for (int t = wv; t < T16 * T16; t += kT / 64) {
      const int xt = t / T16, yt = t - xt * T16;
      const int xa = xt * 16 + lr, yb = yt * 16 + lr;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBM; ++kb) {
        if (kb >= KB) break;
        float a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = kb * 32 + 8 * lg + j;
          a[j] = (k < n && xa < n) ? crr[xa - k + n] : 0.f;
          b[j] = (k < n && yb < n) ? A[yb * rs + k] : 0.f;
        }
        acc = mfma_split3(a, b, acc);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int x = xt * 16 + 4 * lg + v;
        if (x < n && yb < n) B[yb * rs + x] = acc[v];
      }
    }
    __syncthreads();
    float* dst = aout + dout.off[g] + (int64_t)z * nn;
    for (int t = wv; t < T16 * T16; t += kT / 64) {
      const int yt = t / T16, xt = t - yt * T16;
      const int ya = yt * 16 + lr, xb = xt * 16 + lr;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBM; ++kb) {
        if (kb >= KB) break;
        float a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = kb * 32 + 8 * lg + j;
          a[j] = (k < n && ya < n) ? crr[ya - k + n] : 0.f;
          b[j] = (k < n && xb < n) ? B[k * rs + xb] : 0.f;
        }
        acc = mfma_split3(a, b, acc);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int y = yt * 16 + 4 * lg + v;
        if (y < n && xb < n) dst[y * n + xb] = acc[v];
      }
    }
    __syncthreads();
  }
}

// z on blocks of 64 columns: D[z][p] = sum_k C[z][k] Lz[k][p].
template <int NMAX>
__global__ __launch_bounds__(kT) void k_circ_z_bf16(PlanDev P, int cls, DescDev dout, TablesDev tb) {
  constexpr int CB = kColBlock;
  constexpr int KBM = (NMAX + 31) / 32;
  constexpr int SZ = CB + 1;
  __shared__ float Lz[NMAX * SZ];
  __shared__ float crr[2 * NMAX + kRB];
  const int total = P.totals[kQCols * kClasses + cls];
  float* aout = (float*)dout.arena;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lg = lane >> 4;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQCols, tile, g, blk);
    const int n = box_n(dout.bounds, g), nn = n * n;
    const int c0 = blk * CB, width = min(CB, nn - c0);
    const int T16 = (n + 15) >> 4, KB = (n + 31) >> 5;
    load_circ_row(crr, tb, n, tid, kT);
    float* base = aout + dout.off[g];
    stage_rows_to_lds(Lz, SZ, base + c0, nn, n, width, tid, kT);
    __syncthreads();
    for (int t = wv; t < T16 * (CB / 16); t += kT / 64) {
      const int zt = t / (CB / 16), pt = t - zt * (CB / 16);
      const int za = zt * 16 + lr, pb = pt * 16 + lr;
      const bool col = pb < width;
      v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < KBM; ++kb) {
        if (kb >= KB) break;
        float a[8], b[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int k = kb * 32 + 8 * lg + j;
          a[j] = (k < n && za < n) ? crr[za - k + n] : 0.f;
          b[j] = (k < n && col) ? Lz[k * SZ + pb] : 0.f;
        }
        acc = mfma_split3(a, b, acc);
      }
      if (col) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int zz = zt * 16 + 4 * lg + v;
          if (zz < n) base[(int64_t)zz * nn + c0 + pb] = acc[v];
        }
      }
    }
    __syncthreads();
  }
}

// part 0: (x,y) forward / x-then-y circulant; part 1: z; part 2: (y,x) inverse (FP64 only)
template <int NMAX>
void conv_class(const Launch& L, int part, const PlanDev& P, int cls, const DescDev& din, const DescDev& dout,
                const TablesDev& tb, const int64_t* spec_off, void* spec, int mode) {
  const int grid = tile_grid(L, kGridPerCU, cls);
  if (mode == kConvCirculantF32) {
    if (part == 0) hipLaunchKernelGGL(k_circ_xy<NMAX>, dim3(grid), dim3(kT), 0, L.stream, P, cls, din, dout, tb);
    if (part == 1) hipLaunchKernelGGL(k_circ_z<NMAX>, dim3(grid), dim3(kT), 0, L.stream, P, cls, dout, tb);
  } else if (mode == kConvSplitBF16) {
    if (part == 0) hipLaunchKernelGGL(k_circ_xy_bf16<NMAX>, dim3(grid), dim3(kT), 0, L.stream, P, cls, din, dout, tb);
    if (part == 1) hipLaunchKernelGGL(k_circ_z_bf16<NMAX>, dim3(grid), dim3(kT), 0, L.stream, P, cls, dout, tb);
  } else {
    if (part == 0)
      hipLaunchKernelGGL(k_fft_xy_fwd<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, din, tb, spec_off,
                         (double2*)spec);
    if (part == 1)
      hipLaunchKernelGGL(k_fft_z<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, dout, tb, spec_off,
                         (double2*)spec);
    if (part == 2)
      hipLaunchKernelGGL(k_fft_xy_inv<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, dout, tb, spec_off,
                         (const double2*)spec);
  }
}

}  // namespace

void launch_conv(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                 const DescDev& dout, const TablesDev& tb, const int64_t* spec_off, void* spec, int mode) {
  (void)sp;
  conv_class<32>(L, part, P, 0, din, dout, tb, spec_off, spec, mode);
  conv_class<48>(L, part, P, 1, din, dout, tb, spec_off, spec, mode);
  conv_class<64>(L, part, P, 2, din, dout, tb, spec_off, spec, mode);
  conv_class<96>(L, part, P, 3, din, dout, tb, spec_off, spec, mode);
}

}  // namespace gg
