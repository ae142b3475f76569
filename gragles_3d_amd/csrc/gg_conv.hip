// Heat-kernel convolution of every grain box (LSbox::executeConvolution,
// src/box.cpp:414-539, Gaussian branch of convolutionGeneratorMKL :751-767).
//
// FP64 plain form: explicit 3-D DFT on the n^3 box, half spectrum along x,
// spectral factor G, inverse DFT.  Every sum runs over its index from 0
// upward starting at 0.0 with the products grouped as below, and the library
// is built with -ffp-contract=off, so each value is the same IEEE double the
// FP64 CPU restatement produces:
//   forward x (real in):  re = re + v*c ; im = im - v*s
//   forward y, z:         re = re + (a*c + b*s) ; im = im + (b*c - a*s)
//   spectral factor:      re = re*G ; im = im*G
//   inverse z, y:         re = re + (a*c - b*s) ; im = im + (a*s + b*c)
//   inverse x (real out): acc = X0.re ; acc = acc + 2.0*(a*c - b*s), kx=1..(n-1)/2 ;
//                         n even: acc = acc + (x odd ? -Xn/2.re : Xn/2.re)
// with c = cos(2 pi t/n), s = sin(2 pi t/n) host tables and t = (index*k) mod n.
// Three launches per size class: (x,y) forward per z-layer with the layer in
// LDS, z forward + G + z inverse per block of 32 spectral columns, (y,x)
// inverse per z-layer.
//
// FP32 batched form: the Gaussian factor is separable,
// G(kx,ky,kz) = f(kx) f(ky) f(kz), so IDFT . diag(G) . DFT on the box equals
// three 1-D circular convolutions with the real kernel
// c[m] = (1/n) sum_k f(min(k,n-k)) cos(2 pi k m / n) (host long double, stored
// FP32).  Two launches per size class: x then y per z-layer in LDS (vector
// FMAs), then z per block of 64 columns (matrix cores).
#include "gg_internal.h"

namespace gg {
namespace {
constexpr int kGridPerCU = 128;  // workgroups per CU of a convolution launch

constexpr int kT = 256;

template <int NMAX>
__global__ __launch_bounds__(kT) void k_dft_xy_fwd(PlanDev P, int cls, DescDev din, TablesDev tb,
                                                   const int64_t* __restrict__ spec_off, double2* spec) {
  constexpr int R = (NMAX * (NMAX / 2 + 1) + kT - 1) / kT;
  __shared__ double L[NMAX * (NMAX + 2)];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQLayers * kClasses + cls];
  const double* arena = (const double*)din.arena;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(din.bounds, g), n2 = n / 2 + 1, nc = n * n2;
    const int to = tb.tw_off[n];
    for (int i = tid; i < n; i += kT) { tc[i] = tb.twc[to + i]; ts[i] = tb.tws[to + i]; }
    const double* src = arena + din.off[g] + (int64_t)z * n * n;
    stage_to_lds(L, src, n * n, tid, kT);
    __syncthreads();
    double re[R], im[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      double a = 0.0, b = 0.0;
      if (o < nc) {
        const int y = o / n2, kx = o - y * n2;
        const double* v = L + y * n;
        int t = 0;
        for (int x = 0; x < n; ++x) {
          a = a + v[x] * tc[t];
          b = b - v[x] * ts[t];
          t += kx;
          if (t >= n) t -= n;
        }
      }
      re[r] = a;
      im[r] = b;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      if (o < nc) { L[2 * o] = re[r]; L[2 * o + 1] = im[r]; }
    }
    __syncthreads();
    double2* dst = spec + spec_off[g] + (int64_t)z * nc;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      if (o < nc) {
        const int ky = o / n2, kx = o - ky * n2;
        double a = 0.0, b = 0.0;
        int t = 0;
        for (int y = 0; y < n; ++y) {
          const double ar = L[2 * (y * n2 + kx)], ai = L[2 * (y * n2 + kx) + 1];
          a = a + (ar * tc[t] + ai * ts[t]);
          b = b + (ai * tc[t] - ar * ts[t]);
          t += ky;
          if (t >= n) t -= n;
        }
        dst[o] = make_double2(a, b);
      }
    }
    __syncthreads();
  }
}

template <int NMAX>
__global__ __launch_bounds__(kT) void k_dft_z(PlanDev P, int cls, DescDev dref, TablesDev tb,
                                              const int64_t* __restrict__ spec_off, double2* spec) {
  constexpr int CB = kSpecBlock;
  constexpr int R = (NMAX * CB + kT - 1) / kT;
  __shared__ double2 B[NMAX * CB];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQSpecCols * kClasses + cls];
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, blk;
    plan_tile(P, cls, kQSpecCols, tile, g, blk);
    const int n = box_n(dref.bounds, g), n2 = n / 2 + 1, nc = n * n2;
    const int c0 = blk * CB;
    const int to = tb.tw_off[n];
    const double* G = tb.gtab + tb.g_off[n];
    for (int i = tid; i < n; i += kT) { tc[i] = tb.twc[to + i]; ts[i] = tb.tws[to + i]; }
    double2* base = spec + spec_off[g];
    stage_rows_to_lds(B, CB, base + c0, nc, n, min(CB, nc - c0), tid, kT);  // columns past nc are never read
    __syncthreads();
    double fr[R], fi[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      const int kz = o / CB, cb = o - kz * CB, c = c0 + cb;
      double a = 0.0, b = 0.0;
      if (kz < n && c < nc) {
        int t = 0;
        for (int z = 0; z < n; ++z) {
          const double2 v = B[z * CB + cb];
          a = a + (v.x * tc[t] + v.y * ts[t]);
          b = b + (v.y * tc[t] - v.x * ts[t]);
          t += kz;
          if (t >= n) t -= n;
        }
        const int ky = c / n2, kx = c - ky * n2;
        const int kz2 = min(kz, n - kz), ky2 = min(ky, n - ky), kx2 = min(kx, n - kx);
        const double gf = G[kz2 * kz2 + ky2 * ky2 + kx2 * kx2];
        a = a * gf;
        b = b * gf;
      }
      fr[r] = a;
      fi[r] = b;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      if (o < n * CB) B[o] = make_double2(fr[r], fi[r]);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      const int z = o / CB, cb = o - z * CB, c = c0 + cb;
      if (z < n && c < nc) {
        double a = 0.0, b = 0.0;
        int t = 0;
        for (int kz = 0; kz < n; ++kz) {
          const double2 v = B[kz * CB + cb];
          a = a + (v.x * tc[t] - v.y * ts[t]);
          b = b + (v.x * ts[t] + v.y * tc[t]);
          t += z;
          if (t >= n) t -= n;
        }
        base[(int64_t)z * nc + c] = make_double2(a, b);
      }
    }
    __syncthreads();
  }
}

template <int NMAX>
__global__ __launch_bounds__(kT) void k_dft_xy_inv(PlanDev P, int cls, DescDev dout, TablesDev tb,
                                                   const int64_t* __restrict__ spec_off, const double2* spec) {
  constexpr int R = (NMAX * (NMAX / 2 + 1) + kT - 1) / kT;
  constexpr int R3 = (NMAX * NMAX + kT - 1) / kT;
  __shared__ __align__(16) double L[NMAX * (NMAX + 2)];
  __shared__ double tc[NMAX], ts[NMAX];
  const int total = P.totals[kQLayers * kClasses + cls];
  double* arena = (double*)dout.arena;
  const int tid = threadIdx.x;
  for (int tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int g, z;
    plan_tile(P, cls, kQLayers, tile, g, z);
    const int n = box_n(dout.bounds, g), n2 = n / 2 + 1, nc = n * n2;
    const int to = tb.tw_off[n];
    for (int i = tid; i < n; i += kT) { tc[i] = tb.twc[to + i]; ts[i] = tb.tws[to + i]; }
    const double2* src = spec + spec_off[g] + (int64_t)z * nc;
    stage_to_lds((double2*)L, src, nc, tid, kT);
    __syncthreads();
    double re[R], im[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      double a = 0.0, b = 0.0;
      if (o < nc) {
        const int y = o / n2, kx = o - y * n2;
        int t = 0;
        for (int ky = 0; ky < n; ++ky) {
          const double ar = L[2 * (ky * n2 + kx)], ai = L[2 * (ky * n2 + kx) + 1];
          a = a + (ar * tc[t] - ai * ts[t]);
          b = b + (ar * ts[t] + ai * tc[t]);
          t += y;
          if (t >= n) t -= n;
        }
      }
      re[r] = a;
      im[r] = b;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int o = tid + r * kT;
      if (o < nc) { L[2 * o] = re[r]; L[2 * o + 1] = im[r]; }
    }
    __syncthreads();
    double* dst = arena + dout.off[g] + (int64_t)z * n * n;
    const int half = (n - 1) / 2;
#pragma unroll
    for (int r = 0; r < R3; ++r) {
      const int o = tid + r * kT;
      if (o < n * n) {
        const int y = o / n, x = o - y * n;
        const double* row = L + 2 * (y * n2);
        double acc = row[0];
        int t = x;
        for (int kx = 1; kx <= half; ++kx) {
          acc = acc + 2.0 * (row[2 * kx] * tc[t] - row[2 * kx + 1] * ts[t]);
          t += x;
          if (t >= n) t -= n;
        }
        if ((n & 1) == 0) {
          const double a = row[2 * (n / 2)];
          acc = acc + ((x & 1) ? -a : a);
        }
        dst[o] = acc;
      }
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

// part 0: (x,y) forward / x-then-y circulant; part 1: z; part 2: (y,x) inverse (FP64 only)
template <int NMAX>
void conv_class(const Launch& L, int part, const PlanDev& P, int cls, const DescDev& din, const DescDev& dout,
                const TablesDev& tb, const int64_t* spec_off, void* spec, bool fp32) {
  if (fp32) {
    if (part == 0) hipLaunchKernelGGL(k_circ_xy<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, din, dout, tb);
    if (part == 1) hipLaunchKernelGGL(k_circ_z<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, dout, tb);
  } else {
    if (part == 0)
      hipLaunchKernelGGL(k_dft_xy_fwd<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, din, tb, spec_off,
                         (double2*)spec);
    if (part == 1)
      hipLaunchKernelGGL(k_dft_z<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, dout, tb, spec_off,
                         (double2*)spec);
    if (part == 2)
      hipLaunchKernelGGL(k_dft_xy_inv<NMAX>, dim3(tile_grid(L, kGridPerCU, cls)), dim3(kT), 0, L.stream, P, cls, dout, tb, spec_off,
                         (const double2*)spec);
  }
}

}  // namespace

void launch_conv(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                 const DescDev& dout, const TablesDev& tb, const int64_t* spec_off, void* spec, bool fp32) {
  (void)sp;
  conv_class<32>(L, part, P, 0, din, dout, tb, spec_off, spec, fp32);
  conv_class<48>(L, part, P, 1, din, dout, tb, spec_off, spec, fp32);
  conv_class<64>(L, part, P, 2, din, dout, tb, spec_off, spec, fp32);
  conv_class<96>(L, part, P, 3, din, dout, tb, spec_off, spec, fp32);
}

}  // namespace gg
