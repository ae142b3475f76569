// Host-side set-up for the level-set library (include/gg_grainhdl.h): grid and
// time variables, the seeded Voronoi microstructure and the initial neighbour
// lists.  Plain C++ with std::thread workers; no GPU calls.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <atomic>
#include <thread>
#include <vector>

#include "../../include/gg_grainhdl.h"

namespace {

const double kPi = 3.14159265358979323846;

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

void cube_bounds(int32_t* b, int L) {  // resizeToCube, src/dimensionalBuffer.h:139-219
  int grid_size = L - 1;
  int w = b[3] - b[0], hgt = b[4] - b[1], d = b[5] - b[2];
  if (w == hgt && w == d) return;
  int ref = w > hgt ? w : hgt;
  if (ref < d) ref = d;
  if (ref > w) { int diff = ref - w; b[0] -= diff / 2; b[3] += diff / 2 + diff % 2; }
  if (ref > hgt) { int diff = ref - hgt; b[1] -= diff / 2; b[4] += diff / 2 + diff % 2; }
  if (ref > d) { int diff = ref - d; b[2] -= diff / 2; b[5] += diff / 2 + diff % 2; }
  for (int a = 0; a < 3; ++a) {
    if (b[a] < 0) { int delta = -b[a]; b[a + 3] += delta; b[a] += delta; }
    else if (b[a + 3] > grid_size) { int delta = grid_size - b[a + 3]; b[a + 3] += delta; b[a] += delta; }
  }
}

// Run body(i) for i in [0, n) on nthreads host threads, items handed out one
// at a time.
template <class F>
void parallel_for(int n, int nthreads, F body) {
  if (nthreads < 1) {
    nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads < 1) nthreads = 1;
  }
  nthreads = std::min(nthreads, std::max(1, n));
  std::atomic<int> next{0};
  auto worker = [&](int t) {
    for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) body(i, t);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
}

int resolve_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  int n = (int)std::thread::hardware_concurrency();
  return n > 0 ? n : 1;
}

// ---- grain-boundary model of the convolution correction --------------------
struct Quat {
  double q0, q1, q2, q3;
};

// myQuaternion::operator* (src/myQuaternion.cpp:59-71), terms in its order.
Quat qmul(const Quat& a, const Quat& b) {
  Quat r;
  r.q0 = +a.q0 * b.q0 - a.q1 * b.q1 - a.q2 * b.q2 - a.q3 * b.q3;
  r.q1 = +a.q1 * b.q0 + a.q0 * b.q1 - a.q3 * b.q2 + a.q2 * b.q3;
  r.q2 = +a.q2 * b.q0 + a.q3 * b.q1 + a.q0 * b.q2 - a.q1 * b.q3;
  r.q3 = +a.q3 * b.q0 - a.q2 * b.q1 + a.q1 * b.q2 + a.q0 * b.q3;
  return r;
}

// misorientationCubicQxQ (src/myQuaternion.cpp:309-367): the rotation a^-1 b
// taken to the cubic fundamental zone by the largest absolute component over
// its six symmetry-equivalent component sets (24 operators), angle 2 acos.
// computeMisorientation (src/box.cpp:2704-2719) floors it at 1 degree.
double cubic_misorientation(const Quat& a, const Quat& b) {
  const Quat r = qmul(Quat{a.q0, -a.q1, -a.q2, -a.q3}, b);
  const double w = r.q0, x = r.q1, y = r.q2, z = r.q3;
  const double f = 1.0 / sqrt(2.0);
  const double sets[6][4] = {{(w + x) * f, (w - x) * f, (y + z) * f, (y - z) * f},
                             {(w + y) * f, (w - y) * f, (x + z) * f, (x - z) * f},
                             {(w + z) * f, (w - z) * f, (x + y) * f, (x - y) * f},
                             {(w + x + y + z) * 0.5, (w + x - y - z) * 0.5, (w - x + y - z) * 0.5, (w - x - y + z) * 0.5},
                             {(w + x + y - z) * 0.5, (w + x - y + z) * 0.5, (w - x + y + z) * 0.5, (w - x - y - z) * 0.5},
                             {w, x, y, z}};
  double big = 0.0;
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 4; j++)
      if (std::fabs(sets[i][j]) > big) big = std::fabs(sets[i][j]);
  if (big > 1.0) big = (double)(int)big;
  const double angle = 2 * acos(big);
  return angle > 1 * kPi / 180.0 ? angle : 1 * kPi / 180.0;
}

// InterfacialElement::computeMobilityMisori (src/InterfacialElement.cpp:23-35)
double gb_mobility(double theta, bool model) {
  if (!model) return 1.0;
  const double m = 1.0 - (0.99 * exp(-5. * (pow(theta / (15 * kPi / 180), 9.))));
  return m < 0.1 ? 0.1 : m;
}

// InterfacialElement::computeReadShockleyEnergy (src/InterfacialElement.cpp:37-50)
double gb_energy(double theta, bool unique) {
  if (unique) return 1.0;
  const double ref = 15 * kPi / 180;
  if (theta > ref) return 1.0;
  return 1.0 * (theta / ref) * (1.0 - log(theta / ref));
}

double normal_deviate(SplitMix64& r) {  // Box-Muller, u1 in (0, 1]
  const double u1 = 1.0 - r.uniform(), u2 = r.uniform();
  return sqrt(-2.0 * log(u1)) * cos(2 * kPi * u2);
}

}  // namespace

extern "C" {

double gg_misorientation(const double* qa, const double* qb) {
  return cubic_misorientation(Quat{qa[0], qa[1], qa[2], qa[3]}, Quat{qb[0], qb[1], qb[2], qb[3]});
}

int gg_gb_pair_tables(int32_t n_ids, const double* quat, int32_t mobility_model, int32_t unique_energies,
                      double* gam, double* mob, int32_t nthreads) {
  if (n_ids < 1 || !quat || !gam || !mob) return GG_ERR_ARG;
  const int64_t n = n_ids;
  parallel_for(n_ids, nthreads, [&](int a, int) {
    const Quat qa{quat[4 * a], quat[4 * a + 1], quat[4 * a + 2], quat[4 * a + 3]};
    for (int64_t b = 0; b < n; ++b) {
      const double th = cubic_misorientation(qa, Quat{quat[4 * b], quat[4 * b + 1], quat[4 * b + 2], quat[4 * b + 3]});
      gam[a * n + b] = gb_energy(th, unique_energies != 0);
      mob[a * n + b] = gb_mobility(th, mobility_model != 0);
    }
  });
  return GG_OK;
}

void gg_euler_to_quaternion(const double* euler, double* q) {  // euler2Quaternion, src/myQuaternion.cpp:168-191
  const double p1 = euler[0], t = euler[1], p2 = euler[2];
  const double c = cos(t / 2), s = sin(t / 2);
  q[0] = c * cos((p1 + p2) / 2);
  q[1] = s * cos((p1 - p2) / 2);
  q[2] = s * sin((p1 - p2) / 2);
  q[3] = c * sin((p1 + p2) / 2);
}

int gg_random_orientations(int32_t n_grains, uint64_t seed, int32_t reference_abs, double* quat) {
  if (n_grains < 1 || !quat) return GG_ERR_ARG;
  const double wall[3] = {0.0, 35 * (kPi / 180), 0.0};
  gg_euler_to_quaternion(wall, quat);
  SplitMix64 r{seed ^ 0x5851F42D4C957F2Dull};
  for (int g = 1; g <= n_grains; ++g) {
    const double s = r.uniform(), t1 = 2 * kPi * r.uniform(), t2 = 2 * kPi * r.uniform();
    const double s1 = sqrt(1 - s), s2 = sqrt(s);
    double q[4] = {s1 * sin(t1), s1 * cos(t1), s2 * sin(t2), s2 * cos(t2)};
    double nrm = 0;
    for (int i = 0; i < 4; ++i) {
      if (reference_abs) q[i] = std::fabs(q[i]);
      nrm += q[i] * q[i];
    }
    nrm = sqrt(nrm);
    for (int i = 0; i < 4; ++i) quat[4 * g + i] = q[i] / nrm;
  }
  return GG_OK;
}

int gg_lognormal_energies(int32_t n_grains, uint64_t seed, double median, double sigma, double* see) {
  if (n_grains < 1 || !see || !(median > 0) || !(sigma >= 0)) return GG_ERR_ARG;
  see[0] = 0.0;
  SplitMix64 r{seed ^ 0xD1B54A32D192ED03ull};
  for (int g = 1; g <= n_grains; ++g) see[g] = median * exp(sigma * normal_deviate(r));
  return GG_OK;
}

int gg_derive_params(int32_t n_grains, int32_t points_per_grain, int32_t border, double time_step_size,
                     gg_params* out) {
  if (!out || n_grains < 1 || points_per_grain < 1 || border < 0 || !(time_step_size > 0)) return GG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  int rds = (int)(pow((double)n_grains, 1. / 3.0) * pow(kPi * 4 / 3, 1. / 3) / 2 * points_per_grain + 1);
  out->n_grains = n_grains;
  out->grid_blowup = border;
  out->ngridpoints = rds + 2 * border;
  out->h = 1.0 / rds;
  out->delta = border / double(rds);
  out->dt = time_step_size / double(rds * rds);
  out->conv_domain_size = rds - 1;
  out->boundary_grain_tube = 0;
  out->precision = GG_FP64_PLAIN;
  return GG_OK;
}

void gg_resize_to_cube(int32_t* bounds6, int32_t ngridpoints) { cube_bounds(bounds6, ngridpoints); }

int gg_voronoi_seeds(int32_t n_grains, uint64_t seed, double* xyz) {
  if (n_grains < 1 || !xyz) return GG_ERR_ARG;
  SplitMix64 rng{seed};
  for (int i = 0; i < 3 * n_grains; ++i) xyz[i] = rng.uniform();
  return GG_OK;
}

int gg_generate_voronoi(const gg_params* p, uint64_t seed, int32_t* idfield, int32_t* bounds, uint8_t* exists,
                        int32_t nthreads) {
  if (!p || !bounds || !exists) return GG_ERR_ARG;
  const int N = p->n_grains, L = p->ngridpoints, gb = p->grid_blowup;
  const double h = p->h;
  std::vector<double> sx(N), sy(N), sz(N);
  SplitMix64 rng{seed};
  for (int i = 0; i < N; ++i) {
    sx[i] = rng.uniform();
    sy[i] = rng.uniform();
    sz[i] = rng.uniform();
  }
  // bucket grid over [0,1)^3
  int B = std::max(1, (int)std::cbrt((double)N));
  const double cs = 1.0 / B;
  std::vector<int> cstart((size_t)B * B * B + 1, 0), cell(N);
  auto bidx = [&](double v) { int q = (int)(v * B); return q < 0 ? 0 : (q >= B ? B - 1 : q); };
  for (int i = 0; i < N; ++i) {
    cell[i] = (bidx(sz[i]) * B + bidx(sy[i])) * B + bidx(sx[i]);
    cstart[cell[i] + 1]++;
  }
  for (size_t c = 0; c < (size_t)B * B * B; ++c) cstart[c + 1] += cstart[c];
  std::vector<int> items(N), fillp(cstart.begin(), cstart.end() - 1);
  for (int i = 0; i < N; ++i) items[fillp[cell[i]]++] = i;  // ascending id inside a bucket

  const int T = resolve_threads(nthreads);
  std::vector<int32_t> lo((size_t)T * (N + 1) * 3, INT32_MAX), hi((size_t)T * (N + 1) * 3, INT32_MIN);
  parallel_for(L, T, [&](int k, int t) {
    int32_t* mylo = &lo[(size_t)t * (N + 1) * 3];
    int32_t* myhi = &hi[(size_t)t * (N + 1) * 3];
    {
      for (int i = 0; i < L; ++i) {
        for (int j = 0; j < L; ++j) {
          int32_t id = 0;
          if (!(i < gb || j < gb || k < gb || i >= L - 1 - gb || j >= L - 1 - gb || k >= L - 1 - gb)) {
            const double x = double((j - gb) * h), y = double((i - gb) * h), z = double((k - gb) * h);
            const int bx = bidx(x), by = bidx(y), bz = bidx(z);
            double best = 1e300;
            int bid = -1;
            for (int r = 0; r <= B; ++r) {
              // buckets at Chebyshev ring r around (bx,by,bz)
              for (int cz = bz - r; cz <= bz + r; ++cz) {
                if (cz < 0 || cz >= B) continue;
                for (int cy = by - r; cy <= by + r; ++cy) {
                  if (cy < 0 || cy >= B) continue;
                  for (int cx = bx - r; cx <= bx + r; ++cx) {
                    if (cx < 0 || cx >= B) continue;
                    if (std::max(std::abs(cz - bz), std::max(std::abs(cy - by), std::abs(cx - bx))) != r) continue;
                    const int c = (cz * B + cy) * B + cx;
                    for (int q = cstart[c]; q < cstart[c + 1]; ++q) {
                      const int s = items[q];
                      const double dx = x - sx[s], dy = y - sy[s], dz = z - sz[s];
                      const double d2 = dx * dx + dy * dy + dz * dz;
                      if (d2 < best || (d2 == best && s < bid)) {
                        best = d2;
                        bid = s;
                      }
                    }
                  }
                }
              }
              // every seed in ring r+1 or beyond lies at least r*cs away
              const double rr = r * cs;
              if (bid >= 0 && best < rr * rr) break;
            }
            id = bid + 1;
            int32_t* l = mylo + 3 * id;
            int32_t* u = myhi + 3 * id;
            if (j < l[0]) l[0] = j;
            if (j > u[0]) u[0] = j;
            if (i < l[1]) l[1] = i;
            if (i > u[1]) u[1] = i;
            if (k < l[2]) l[2] = k;
            if (k > u[2]) u[2] = k;
          }
          if (idfield) idfield[((size_t)k * L + i) * L + j] = id;
        }
      }
    }
  });
  std::memset(bounds, 0, sizeof(int32_t) * 6 * (N + 1));
  exists[0] = 0;
  for (int g = 1; g <= N; ++g) {
    int32_t l[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, u[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int t = 0; t < T; ++t)
      for (int a = 0; a < 3; ++a) {
        l[a] = std::min(l[a], lo[((size_t)t * (N + 1) + g) * 3 + a]);
        u[a] = std::max(u[a], hi[((size_t)t * (N + 1) + g) * 3 + a]);
      }
    if (l[0] == INT32_MAX) {
      exists[g] = 0;
      continue;
    }
    exists[g] = 1;
    for (int a = 0; a < 3; ++a) {
      bounds[6 * g + a] = std::max(0, l[a] - gb);
      bounds[6 * g + 3 + a] = std::min(L, u[a] + 1 + gb);
    }
  }
  return GG_OK;
}

int gg_bimodal_seeds(int32_t n_grains, uint64_t seed, double frac_large, double size_ratio, double sigma,
                     double* xyz, double* radius) {
  if (n_grains < 1 || !xyz || !radius || !(frac_large >= 0 && frac_large <= 1) || !(size_ratio > 0) ||
      !(sigma >= 0))
    return GG_ERR_ARG;
  const int N = n_grains;
  SplitMix64 rng{seed};
  // radii: the mode (1 or size_ratio) times a lognormal factor exp(sigma z),
  // z standard normal by Box-Muller from two uniforms
  std::vector<double> r(N);
  double vol = 0;
  for (int i = 0; i < N; ++i) {
    const bool large = rng.uniform() < frac_large;
    const double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
    const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    r[i] = (large ? size_ratio : 1.0) * std::exp(sigma * z);
    vol += r[i] * r[i] * r[i];
  }
  // scale: the balls' volumes add up to the unit cube's
  const double s = std::cbrt(1.0 / (4.0 / 3.0 * kPi * vol));
  double rmax = 0;
  for (int i = 0; i < N; ++i) {
    r[i] *= s;
    rmax = std::max(rmax, r[i]);
  }
  // random sequential addition, largest first: a seed is kept if its ball of
  // radius alpha * r_i misses the ball every earlier seed kept (alpha shrinks
  // for a grain that finds no room in kTries attempts)
  std::vector<int> ord(N);
  for (int i = 0; i < N; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return r[a] > r[b]; });
  constexpr double kAlpha = 0.85;
  constexpr int kTries = 200;
  const double cs = kAlpha * 2.0 * rmax;  // a bucket spans the largest exclusion distance
  const int B = std::max(1, std::min(256, (int)(1.0 / cs)));
  std::vector<std::vector<int>> bucket((size_t)B * B * B);
  auto bi = [&](double v) { int q = (int)(v * B); return q < 0 ? 0 : (q >= B ? B - 1 : q); };
  for (int i : ord) {
    double alpha = kAlpha, px = 0, py = 0, pz = 0;
    for (int attempt = 0;; ++attempt) {
      if (attempt > 0 && attempt % kTries == 0) alpha *= 0.9;
      px = rng.uniform();
      py = rng.uniform();
      pz = rng.uniform();
      const int bx = bi(px), by = bi(py), bz = bi(pz);
      bool ok = true;
      for (int cz = std::max(0, bz - 1); ok && cz <= std::min(B - 1, bz + 1); ++cz)
        for (int cy = std::max(0, by - 1); ok && cy <= std::min(B - 1, by + 1); ++cy)
          for (int cx = std::max(0, bx - 1); ok && cx <= std::min(B - 1, bx + 1); ++cx)
            for (int j : bucket[((size_t)cz * B + cy) * B + cx]) {
              const double dx = px - xyz[3 * j], dy = py - xyz[3 * j + 1], dz = pz - xyz[3 * j + 2];
              const double m = alpha * r[i] + radius[j];  // the ball j keeps
              if (dx * dx + dy * dy + dz * dz < m * m) {
                ok = false;
                break;
              }
            }
      if (ok) break;
    }
    xyz[3 * i] = px;
    xyz[3 * i + 1] = py;
    xyz[3 * i + 2] = pz;
    bucket[((size_t)bi(pz) * B + bi(py)) * B + bi(px)].push_back(i);
    radius[i] = alpha * r[i];
  }
  return GG_OK;
}

int gg_generate_power(const gg_params* p, const double* xyz, const double* radius, int32_t* idfield,
                      int32_t* bounds, uint8_t* exists, int32_t nthreads) {
  if (!p || !xyz || !radius || !bounds || !exists) return GG_ERR_ARG;
  const int N = p->n_grains, L = p->ngridpoints, gb = p->grid_blowup;
  const double h = p->h;
  std::vector<double> w(N);
  double wmax = 0;
  for (int i = 0; i < N; ++i) {
    w[i] = radius[i] * radius[i];
    wmax = std::max(wmax, w[i]);
  }
  // seed buckets over [0,1)^3 with each bucket's largest weight
  const int B = std::max(1, (int)std::cbrt((double)N));
  const double cs = 1.0 / B;
  auto bidx = [&](double v) { int q = (int)(v * B); return q < 0 ? 0 : (q >= B ? B - 1 : q); };
  std::vector<int> cstart((size_t)B * B * B + 1, 0), cell(N);
  for (int i = 0; i < N; ++i) {
    cell[i] = (bidx(xyz[3 * i + 2]) * B + bidx(xyz[3 * i + 1])) * B + bidx(xyz[3 * i]);
    cstart[cell[i] + 1]++;
  }
  for (size_t c = 0; c < (size_t)B * B * B; ++c) cstart[c + 1] += cstart[c];
  std::vector<int> items(N), fillp(cstart.begin(), cstart.end() - 1);
  for (int i = 0; i < N; ++i) items[fillp[cell[i]]++] = i;
  std::vector<double> bw((size_t)B * B * B, 0.0);
  for (int i = 0; i < N; ++i) bw[cell[i]] = std::max(bw[cell[i]], w[i]);
  // voxel blocks of 8^3 over the interior: the seeds that can own a voxel of
  // the block are those whose smallest power distance over the block does not
  // exceed the smallest largest-power-distance U of any seed (exact: another
  // seed is then strictly better everywhere in the block)
  constexpr int kBlk = 8;
  const int lo = gb, hi = L - 1 - gb;  // interior voxels: lo <= i, j, k < hi
  const int nb = hi > lo ? (hi - lo + kBlk - 1) / kBlk : 0;
  const int T = resolve_threads(nthreads);
  std::vector<int32_t> blo((size_t)T * (N + 1) * 3, INT32_MAX), bhi((size_t)T * (N + 1) * 3, INT32_MIN);
  if (idfield) std::memset(idfield, 0, sizeof(int32_t) * (size_t)L * L * L);
  parallel_for(nb * nb * nb, T, [&](int blk, int t) {
    const int kb = blk / (nb * nb), ib = (blk / nb) % nb, jb = blk % nb;
    const int k0 = lo + kb * kBlk, i0 = lo + ib * kBlk, j0 = lo + jb * kBlk;
    const int k1 = std::min(hi, k0 + kBlk), i1 = std::min(hi, i0 + kBlk), j1 = std::min(hi, j0 + kBlk);
    // the block's box in domain coordinates (x from j, y from i, z from k)
    const double ax = (j0 - gb) * h, bx_ = (j1 - 1 - gb) * h, ay = (i0 - gb) * h, by_ = (i1 - 1 - gb) * h,
                 az = (k0 - gb) * h, bz_ = (k1 - 1 - gb) * h;
    auto lbound = [&](double px, double py, double pz) {  // squared distance, box to point
      const double dx = px < ax ? ax - px : (px > bx_ ? px - bx_ : 0.0);
      const double dy = py < ay ? ay - py : (py > by_ ? py - by_ : 0.0);
      const double dz = pz < az ? az - pz : (pz > bz_ ? pz - bz_ : 0.0);
      return dx * dx + dy * dy + dz * dz;
    };
    auto ubound = [&](double px, double py, double pz) {  // largest squared distance over the box
      const double dx = std::max(std::fabs(px - ax), std::fabs(px - bx_));
      const double dy = std::max(std::fabs(py - ay), std::fabs(py - by_));
      const double dz = std::max(std::fabs(pz - az), std::fabs(pz - bz_));
      return dx * dx + dy * dy + dz * dz;
    };
    const int cx0 = bidx(ax), cx1 = bidx(bx_), cy0 = bidx(ay), cy1 = bidx(by_), cz0 = bidx(az), cz1 = bidx(bz_);
    double U = 1e300;
    std::vector<int> cand;
    std::vector<double> candlb;
    for (int r = 0; r <= B; ++r) {
      // buckets at Chebyshev distance r from the block's bucket range
      for (int cz = cz0 - r; cz <= cz1 + r; ++cz) {
        if (cz < 0 || cz >= B) continue;
        for (int cy = cy0 - r; cy <= cy1 + r; ++cy) {
          if (cy < 0 || cy >= B) continue;
          for (int cx = cx0 - r; cx <= cx1 + r; ++cx) {
            if (cx < 0 || cx >= B) continue;
            const int ring = std::max(std::max(std::max(cz0 - cz, cz - cz1), std::max(cy0 - cy, cy - cy1)),
                                      std::max(cx0 - cx, cx - cx1));
            if (ring != r) continue;
            const int c = (cz * B + cy) * B + cx;
            for (int q = cstart[c]; q < cstart[c + 1]; ++q) {
              const int s = items[q];
              const double sx = xyz[3 * s], sy = xyz[3 * s + 1], sz = xyz[3 * s + 2];
              const double lb = lbound(sx, sy, sz) - w[s];
              if (lb > U) continue;
              U = std::min(U, ubound(sx, sy, sz) - w[s]);
              cand.push_back(s);
              candlb.push_back(lb);
            }
          }
        }
      }
      // a seed in a bucket at ring > r lies more than r*cs from the block
      const double dmin = r * cs;
      if (U < 1e300 && dmin * dmin - wmax > U) break;
    }
    // keep the candidates that can win, in id order (ties go to the lowest id)
    std::vector<int> keep;
    const double slack = 1e-12 * (std::fabs(U) + 1.0);  // rounding of the two bounds
    for (size_t q = 0; q < cand.size(); ++q)
      if (candlb[q] <= U + slack) keep.push_back(cand[q]);
    std::sort(keep.begin(), keep.end());
    int32_t* mylo = &blo[(size_t)t * (N + 1) * 3];
    int32_t* myhi = &bhi[(size_t)t * (N + 1) * 3];
    for (int k = k0; k < k1; ++k)
      for (int i = i0; i < i1; ++i)
        for (int j = j0; j < j1; ++j) {
          const double x = double((j - gb) * h), y = double((i - gb) * h), z = double((k - gb) * h);
          double best = 1e300;
          int bid = -1;
          for (int s : keep) {
            const double dx = x - xyz[3 * s], dy = y - xyz[3 * s + 1], dz = z - xyz[3 * s + 2];
            const double d2 = dx * dx + dy * dy + dz * dz - w[s];
            if (d2 < best) {
              best = d2;
              bid = s;
            }
          }
          const int32_t id = bid + 1;
          int32_t* l = mylo + 3 * id;
          int32_t* u = myhi + 3 * id;
          if (j < l[0]) l[0] = j;
          if (j > u[0]) u[0] = j;
          if (i < l[1]) l[1] = i;
          if (i > u[1]) u[1] = i;
          if (k < l[2]) l[2] = k;
          if (k > u[2]) u[2] = k;
          if (idfield) idfield[((size_t)k * L + i) * L + j] = id;
        }
  });
  std::memset(bounds, 0, sizeof(int32_t) * 6 * (N + 1));
  exists[0] = 0;
  for (int g = 1; g <= N; ++g) {
    int32_t l[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, u[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
    for (int t = 0; t < T; ++t)
      for (int a = 0; a < 3; ++a) {
        l[a] = std::min(l[a], blo[((size_t)t * (N + 1) + g) * 3 + a]);
        u[a] = std::max(u[a], bhi[((size_t)t * (N + 1) + g) * 3 + a]);
      }
    if (l[0] == INT32_MAX) {
      exists[g] = 0;
      continue;
    }
    exists[g] = 1;
    for (int a = 0; a < 3; ++a) {
      bounds[6 * g + a] = std::max(0, l[a] - gb);
      bounds[6 * g + 3 + a] = std::min(L, u[a] + 1 + gb);
    }
  }
  return GG_OK;
}

int gg_initial_neighbour_lists(const gg_params* p, const int32_t* bounds, const uint8_t* exists, int32_t* offsets,
                               int32_t* ids, int64_t cap, int64_t* n) {
  if (!p || !bounds || !exists || !offsets || !n) return GG_ERR_ARG;
  const int N = p->n_grains, L = p->ngridpoints;
  std::vector<int32_t> cb(6 * (N + 1), 0);
  std::vector<int> order;
  int maxside = 1;
  for (int g = 1; g <= N; ++g) {
    if (!exists[g]) continue;
    for (int a = 0; a < 6; ++a) cb[6 * g + a] = bounds[6 * g + a];
    cube_bounds(&cb[6 * g], L);
    maxside = std::max(maxside, cb[6 * g + 3] - cb[6 * g]);
    order.push_back(g);
  }
  std::sort(order.begin(), order.end(), [&](int a, int b) { return cb[6 * a] < cb[6 * b] || (cb[6 * a] == cb[6 * b] && a < b); });
  std::vector<std::vector<int32_t>> nb(N + 1);
  parallel_for((int)order.size(), 0, [&](int oi, int) {
    const int g = order[oi];
    const int32_t* a = &cb[6 * g];
    // candidates h with h.x0 in [a.x0 - maxside, a.x1]
    auto first = std::lower_bound(order.begin(), order.end(), a[0] - maxside,
                                  [&](int q, int v) { return cb[6 * q] < v; });
    std::vector<int32_t>& out = nb[g];
    for (auto it = first; it != order.end() && cb[6 * *it] <= a[3]; ++it) {
      const int hgr = *it;
      if (hgr == g) continue;
      const int32_t* b = &cb[6 * hgr];
      // inclusive overlap, as the R-tree search and checkIntersection test it
      if (a[0] > b[3] || a[3] < b[0] || a[1] > b[4] || a[4] < b[1] || a[2] > b[5] || a[5] < b[2]) continue;
      out.push_back(hgr);
    }
    std::sort(out.begin(), out.end());
  });
  int64_t total = 0;
  offsets[0] = 0;
  offsets[1] = 0;
  for (int g = 1; g <= N; ++g) {
    offsets[g + 1] = (int32_t)(offsets[g] + nb[g].size());
    total += (int64_t)nb[g].size();
  }
  *n = total;
  if (total > cap || !ids) return total > cap ? GG_ERR_CAPACITY : GG_OK;
  for (int g = 1; g <= N; ++g) std::copy(nb[g].begin(), nb[g].end(), ids + offsets[g]);
  return GG_OK;
}

}  // extern "C"
