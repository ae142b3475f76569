// Context management and the C ABI of libgg_levelset (include/gg_levelset.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gg_grainhdl.h"
#include "../../include/gg_levelset.h"
#include "gg_internal.h"

using namespace gg;

// Scratch of gg_init_distance_noalloc: the ID field of a run of grid layers,
// staged from host memory.  4 Mi int32 entries (16 MiB) per GPU, held until the
// process ends; a grid needs ngridpoints^2 <= kNoAllocScratch (ngridpoints <=
// 2048).  Host code uses the device address from hipGetSymbolAddress, taken
// once per device when the first context on it is created.
constexpr int64_t kNoAllocScratch = int64_t(1) << 22;
static __device__ __align__(256) int32_t scratch[kNoAllocScratch];
constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu[kScratchDevices];          // one call at a time per GPU
static int32_t* g_scratch_ptr[kScratchDevices] = {};      // device address of scratch, per GPU

struct gg_ctx {
  gg_params p{};
  SimParams sp{};
  bool fp32 = false;
  bool conv_bf16 = false;  // FP32 path: split-BF16 circulant on the matrix cores
  size_t esz = 8;
  hipStream_t stream = nullptr;
  std::string err;
  int grid = 2048;

  // per grain
  int32_t* d_state = nullptr;
  int32_t* d_newb = nullptr;
  int32_t* d_cnb = nullptr;  // resizeGrid's bounds of the last grid coarsening (allocated on first use)
  int32_t* d_vorder = nullptr;  // volume-sorted work order of the plans (allocated on first use)
  uint32_t* d_cont = nullptr;  // voxel container (allocated on first use)
  int64_t cont_cap = 0;
  TopoDev topo{nullptr, nullptr, nullptr, 0, nullptr, nullptr};  // hull topology tables (allocated when first enabled)
  int topo_user = 0;  // topology asked for through gg_enable_hull_topology (the correction needs it too)
  // convolution correction (gg_set_physics)
  bool corr_on = false;
  double tl_drag = 0.0;
  CorrDev cd{};
  int32_t* d_exflag = nullptr;
  double* d_volume = nullptr;
  double* d_area = nullptr;
  // distance roles
  DescDev D[2]{};
  int sel = 0;  // D[sel] is the input role
  void* arena[2] = {nullptr, nullptr};
  int64_t arena_cap = 0;
  // IDLocal
  IdlDev IL[2]{};
  int il_cur = 0;  // IL[il_cur] holds the latest IDLocal content
  // lists
  ListDev LD{};
  // plan
  PlanDev P{};   // plan of the current boxes (input role)
  PlanDev P2{};  // plan of the output boxes, from the level set's first part to the end of the redistancing
  bool p2_valid = false;  // P2 holds the plan of the current output boxes (the redistancing runs over it)
  int cmp_variant[kClasses] = {kCmpDefault, kCmpDefault, kCmpDefault, kCmpDefault};  // comparison tile shape per class
  int cmp_order = GG_CMP_ORDER_GRAIN;  // comparison tiles per class in grain order, or all in spatial order
  int cmp_space_shape = 0;             // tile shape of the spatially ordered launch
  bool cmp_prebuilt = false;           // neighbour table and spatial work list already enqueued (side stream)
  CmpSpaceDev W{};                     // its work list (allocated with the arenas)
  int64_t* d_spec_off = nullptr;
  void* d_spec = nullptr;
  int64_t spec_cap = 0;
  double* d_partials = nullptr;
  // device scalars
  int32_t* d_err = nullptr;
  int32_t* d_step = nullptr;
  int32_t* d_hist = nullptr;
  int32_t hist_cap = 1 << 16;
  int64_t* d_voxels = nullptr;
  // tables
  TablesDev TB{};
  bool boxes_set = false;
  bool lists_set = false;
  // graph of two timesteps
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_sel = -1, graph_il = -1;
  const int32_t* graph_plan = nullptr;  // c->P.grains when the graph was captured
  // stage timing
  bool timing = false;
  // second stream for the list build, which runs beside the convolution
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev[12] = {};
  double seg_ms[GG_NUM_SEGMENTS] = {};
  int64_t seg_n[GG_NUM_SEGMENTS] = {};
  std::vector<std::pair<int, int>> pending;  // (stage, event pair index)
  std::vector<hipEvent_t> evpool;
  int evused = 0;

  std::vector<void*> allocs;
};

namespace {

int fail(gg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, call)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail((c), GG_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
  } while (0)

template <class T>
int dalloc(gg_ctx* c, T** ptr, size_t count) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(count * sizeof(T), 256);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) return fail(c, GG_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.push_back(p);
  *ptr = (T*)p;
  return GG_OK;
}

// Frees a dalloc'd buffer before the context ends.
void dfree(gg_ctx* c, const void* ptr) {
  auto it = std::find(c->allocs.begin(), c->allocs.end(), ptr);
  if (it == c->allocs.end()) return;
  hipFree(*it);
  c->allocs.erase(it);
}

const double kPi = 3.14159265358979323846;

Launch launch_of(gg_ctx* c) { return Launch{c->stream, c->grid, c->grid / 8}; }

int check_device_error(gg_ctx* c) {
  int32_t e = 0;
  HIPCHK(c, hipMemcpy(&e, c->d_err, sizeof(e), hipMemcpyDeviceToHost));
  if (e) {
    std::string m = "device capacity error:";
    if (e & kErrArena) m += " distance arena full;";
    if (e & kErrList) m += " comparison list capacity exceeded;";
    if (e & kErrBoxSize) m += " box side above 96 or box not cubic;";
    if (e & kErrSpec) m += " spectral scratch full;";
    if (e & kErrEmptyIO) m += " empty in/out intersection in redistancing;";
    if (e & kErrHistory) m += " count history full;";
    if (e & kErrTopo) m += " hull topology table full;";
    if (e & kErrTri) m += " correction barycenter pool full;";
    if (e & kErrCmpWork) m += " comparison work list full;";
    return fail(c, GG_ERR_CAPACITY, m);
  }
  return GG_OK;
}

// The C library's exp/cos/sin, called through volatile pointers so the
// compiler cannot fold, fuse (sincos) or vectorise the calls: the FP64 tables
// then carry exactly the bits of the host C library, as the CPU restatement's
// do on the same host.
double (*volatile libm_exp)(double) = ::exp;
double (*volatile libm_cos)(double) = ::cos;
double (*volatile libm_sin)(double) = ::sin;

// host copies of the per-n tables
int build_tables(gg_ctx* c) {
  std::vector<int32_t> tw_off(kMaxN + 1, 0), g_off(kMaxN + 1, 0);
  std::vector<double> twc, tws, gt;
  std::vector<float> circ;
  const int nn = c->p.conv_domain_size;
  const double dt = c->p.dt;
  for (int n = 1; n <= kMaxN; ++n) {
    tw_off[n] = (int32_t)twc.size();
    for (int t = 0; t < n; ++t) {
      twc.push_back(libm_cos(2.0 * kPi * t / n));
      tws.push_back(libm_sin(2.0 * kPi * t / n));
    }
    // spectral factor, the reference's expression evaluated left to right (src/box.cpp:760-761)
    g_off[n] = (int32_t)gt.size();
    const double nsq = nn * nn;
    const double local_nsq = n * n;
    const int h2 = n / 2;
    for (int q = 0; q <= 3 * h2 * h2; ++q)
      gt.push_back(libm_exp(-sqrt(2 / kPi) * q * 4.0 * dt * nsq / local_nsq * kPi * kPi) / (local_nsq * n));
    // one-axis circulant row: c[m] = (1/n) sum_k f(min(k,n-k)) cos(2 pi k m / n)
    std::vector<double> f(n);
    for (int k = 0; k < n; ++k) {
      int k2 = std::min(k, n - k);
      f[k] = exp(-sqrt(2 / kPi) * (k2 * k2) * 4.0 * dt * nsq / local_nsq * kPi * kPi);
    }
    // cos(2 pi t / n) takes n values; (k m) mod n picks one
    std::vector<long double> ct(n);
    for (int t = 0; t < n; ++t) ct[t] = cosl(2.0L * 3.141592653589793238462643383279L * t / n);
    for (int m = 0; m < n; ++m) {
      long double acc = 0;
      for (int k = 0; k < n; ++k) acc += (long double)f[k] * ct[(k * m) % n];
      circ.push_back((float)(acc / n));
    }
  }
  // FFT radices per n: 4 while 4 divides, then 2, then the odd prime factors ascending
  std::vector<int32_t> fplan((kMaxN + 1) * 8, 0);
  for (int n = 1; n <= kMaxN; ++n) {
    int m = n, s = 0;
    while (m % 4 == 0) { fplan[n * 8 + s++] = 4; m /= 4; }
    if (m % 2 == 0) { fplan[n * 8 + s++] = 2; m /= 2; }
    for (int q = 3; m > 1; q += 2)
      while (m % q == 0) { fplan[n * 8 + s++] = q; m /= q; }
  }
  int32_t *d_two, *d_go, *d_fp;
  double *d_twc, *d_tws, *d_gt;
  float* d_circ;
  int rc;
  if ((rc = dalloc(c, &d_two, tw_off.size())) || (rc = dalloc(c, &d_go, g_off.size())) ||
      (rc = dalloc(c, &d_twc, twc.size())) || (rc = dalloc(c, &d_tws, tws.size())) ||
      (rc = dalloc(c, &d_gt, gt.size())) || (rc = dalloc(c, &d_circ, circ.size())) ||
      (rc = dalloc(c, &d_fp, fplan.size())))
    return rc;
  HIPCHK(c, hipMemcpy(d_two, tw_off.data(), tw_off.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_go, g_off.data(), g_off.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_twc, twc.data(), twc.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_tws, tws.data(), tws.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_gt, gt.data(), gt.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_circ, circ.data(), circ.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(d_fp, fplan.data(), fplan.size() * 4, hipMemcpyHostToDevice));
  c->TB = TablesDev{d_two, d_twc, d_tws, d_go, d_gt, d_circ, d_fp};
  return GG_OK;
}

void cube_bounds_host(int32_t* b, int L) {
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

DescDev din(gg_ctx* c) { return c->D[c->sel]; }
DescDev dout(gg_ctx* c) { return c->D[c->sel ^ 1]; }

void plan_into(gg_ctx* c, const DescDev& d, const PlanDev& P) {
  launch_plan(launch_of(c), c->sp, c->d_state, d.bounds, d.off, c->d_spec_off, P, c->arena_cap, c->spec_cap,
              c->d_err, c->d_voxels);
}
void plan_layout(gg_ctx* c, const DescDev& d) { plan_into(c, d, c->P); }

// stage timing helpers
hipEvent_t take_event(gg_ctx* c) {
  if (c->evused >= (int)c->evpool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->evpool.push_back(e);
  }
  return c->evpool[c->evused++];
}

// Event-timed segments, one per kernel (or per kernel with one launch per size class).
enum Seg {
  kSegConvA, kSegConvB, kSegConvC, kSegIdl, kSegLists, kSegCompare, kSegHull, kSegFinalize, kSegPlan,
  kSegRedistXY, kSegRedistZ, kSegPost, kSegCount, kSegCompare48, kSegCompare64, kSegCompare96, kSegResample,
  kSegContainer, kSegCorrect, kSegElements, kNumSeg
};
static_assert(kNumSeg == GG_NUM_SEGMENTS, "segment table of the C ABI");
static_assert(kCmpVariants == GG_CMP_VARIANTS, "comparison variants of the C ABI");

struct SegTimer {
  gg_ctx* c;
  int seg;
  hipEvent_t a = nullptr;
  SegTimer(gg_ctx* ctx, int s) : c(ctx), seg(s) {
    if (c->timing) {
      a = take_event(c);
      if (a) hipEventRecord(a, c->stream);
    }
  }
  ~SegTimer() {
    if (c->timing && a) {
      hipEvent_t b = take_event(c);
      if (b) {
        hipEventRecord(b, c->stream);
        c->pending.push_back({seg, c->evused - 2});
      }
    }
  }
};

void collect_timing(gg_ctx* c) {
  for (auto& pr : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, c->evpool[pr.second], c->evpool[pr.second + 1]) == hipSuccess) {
      c->seg_ms[pr.first] += ms;
      c->seg_n[pr.first] += 1;
    }
  }
  c->pending.clear();
  c->evused = 0;
}

// Launch arguments of the correction kernels: the time step and the triple-line
// mobility term follow the grid (grid coarsening changes both).
CorrDev corr_args(gg_ctx* c) {
  CorrDev d = c->cd;
  d.dt = c->p.dt;
  const double ds = 1.0 / (double)c->p.conv_domain_size;  // grainhdl::get_ds()
  const double dsl = 3 * sqrt(3 * ds * ds);
  d.tl_inv = 1. / (dsl * c->tl_drag);
  return d;
}

// The interfacial elements of the hull just computed (after the level set's
// last part, or a coarsening's extractContour).
void stage_elements(gg_ctx* c) {
  if (!c->corr_on) return;
  SegTimer t(c, kSegElements);
  launch_elements(launch_of(c), c->sp, c->d_state, c->topo, corr_args(c), c->d_err);
}

// A captured step graph holds kernel arguments and shapes of the moment.
void drop_graph(gg_ctx* c) {
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

int require_boxes(gg_ctx* c) {
  if (!c->boxes_set) return fail(c, GG_ERR_STATE, "gg_set_boxes has not been called");
  return GG_OK;
}

int conv_mode(const gg_ctx* c) {
  if (!c->fp32) return kConvFFT64;
  return c->conv_bf16 ? kConvSplitBF16 : kConvCirculantF32;
}

// ---- stages (each enqueues its kernels on the context stream) -------------
void stage_convolution(gg_ctx* c) {
  Launch L = launch_of(c);
  { SegTimer t(c, kSegConvA); launch_conv(L, 0, c->sp, c->P, din(c), dout(c), c->TB, c->d_spec_off, c->d_spec, conv_mode(c)); }
  { SegTimer t(c, kSegConvB); launch_conv(L, 1, c->sp, c->P, din(c), dout(c), c->TB, c->d_spec_off, c->d_spec, conv_mode(c)); }
  if (!c->fp32) {
    SegTimer t(c, kSegConvC);
    launch_conv(L, 2, c->sp, c->P, din(c), dout(c), c->TB, c->d_spec_off, c->d_spec, conv_mode(c));
  }
  { SegTimer t(c, kSegIdl); launch_idl_resize(L, c->sp, c->d_state, dout(c), c->IL[c->il_cur ^ 1]); }
  if (c->corr_on) {
    // IL[il_cur] holds the last comparison's IDLocal, read as its resize to
    // the output bounds leaves it
    SegTimer t(c, kSegCorrect);
    launch_correction(L, c->sp, c->P, c->d_state, din(c), dout(c), c->IL[c->il_cur], c->topo, corr_args(c),
                      c->fp32);
  }
}

void stage_lists(gg_ctx* c) {
  SegTimer t(c, kSegLists);
  launch_lists(launch_of(c), c->sp, c->d_state, din(c), dout(c), c->LD, c->d_newb, c->d_err);
}

void stage_comparison(gg_ctx* c) {
  static const int seg[kClasses] = {kSegCompare, kSegCompare48, kSegCompare64, kSegCompare96};
  if (c->cmp_order == GG_CMP_ORDER_SPACE) {
    SegTimer t(c, kSegCompare);
    if (!c->cmp_prebuilt) launch_comparison_table(launch_of(c), c->sp, c->d_state, din(c), c->LD);
    launch_comparison_space(launch_of(c), c->cmp_space_shape, c->sp, c->P, din(c), dout(c), c->LD,
                            c->IL[c->il_cur], c->IL[c->il_cur ^ 1], c->d_newb, c->W, c->d_err, c->fp32,
                            c->cmp_prebuilt ? kSpaceRun : kSpaceAll);
    c->cmp_prebuilt = false;
  } else if (c->timing) {
    for (int cls = 0; cls < kClasses; ++cls) {
      SegTimer t(c, seg[cls]);
      // the first class's segment includes building the neighbour table
      if (cls == 0) launch_comparison_table(launch_of(c), c->sp, c->d_state, din(c), c->LD);
      launch_comparison(launch_of(c), cls, c->cmp_variant[cls], c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur],
                        c->IL[c->il_cur ^ 1], c->d_newb, c->fp32);
    }
  } else {
    // The size classes touch disjoint grains (their bounding-box atomics and
    // IDLocal rows included), so the n <= 32 class runs on the side stream
    // beside the others; each launch's tail then overlaps the other's work.
    launch_comparison_table(launch_of(c), c->sp, c->d_state, din(c), c->LD);
    hipEventRecord(c->ev_fork, c->stream);
    hipStreamWaitEvent(c->side, c->ev_fork, 0);
    Launch S = launch_of(c);
    S.stream = c->side;
    launch_comparison(S, 0, c->cmp_variant[0], c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur], c->IL[c->il_cur ^ 1], c->d_newb,
                      c->fp32);
    hipEventRecord(c->ev_join, c->side);
    for (int cls = 1; cls < kClasses; ++cls)
      launch_comparison(launch_of(c), cls, c->cmp_variant[cls], c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur],
                        c->IL[c->il_cur ^ 1], c->d_newb, c->fp32);
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
  }
  c->il_cur ^= 1;
}

// Runs launch(L, c0, c1) over the size classes: outside stage timing the
// n <= 32 class goes on the side stream beside the others (classes touch
// disjoint grains), forked from and joined back to the main stream.
template <class F>
void over_classes(gg_ctx* c, F launch) {
  Launch L = launch_of(c);
  if (c->timing) {
    launch(L, 0, kClasses);
    return;
  }
  hipEventRecord(c->ev_fork, c->stream);
  hipStreamWaitEvent(c->side, c->ev_fork, 0);
  Launch S = L;
  S.stream = c->side;
  launch(S, 0, 1);
  hipEventRecord(c->ev_join, c->side);
  launch(L, 1, kClasses);
  hipStreamWaitEvent(c->stream, c->ev_join, 0);
}

// Level set, first part: the output boxes (they follow from the
// set-comparison bounds alone) and their plan, made before the hull so that the
// redistancing need not wait for it.  Grains the hull then finds vanished keep
// their planned space unused.
void stage_output_plan(gg_ctx* c) {
  Launch L = launch_of(c);
  {
    SegTimer t(c, kSegFinalize);
    launch_finalize(L, c->sp, c->P, c->d_state, c->LD, c->d_exflag, c->d_partials, c->d_newb, din(c), dout(c),
                    c->d_volume, c->d_area, kFinBounds);
  }
  { SegTimer t(c, kSegPlan); plan_into(c, dout(c), c->P2); }
  c->p2_valid = true;
}

void stage_hull(gg_ctx* c) {
  SegTimer t(c, kSegHull);
  if (c->topo.on) launch_topo_clear(launch_of(c), c->sp, c->topo);
  over_classes(c, [&](const Launch& LL, int c0, int c1) {
    launch_hull(LL, c->sp, c->P, din(c), c->d_exflag, c->d_partials, c->IL[c->il_cur], c->topo, c->d_err, c->fp32,
                c0, c1);
  });
}

// Level set, last part: deletion, list flip, existence, volume and area.
void stage_finalize_state(gg_ctx* c) {
  SegTimer t(c, kSegFinalize);
  launch_finalize(launch_of(c), c->sp, c->P, c->d_state, c->LD, c->d_exflag, c->d_partials, c->d_newb, din(c),
                  dout(c), c->d_volume, c->d_area, kFinState);
}

void stage_level_set(gg_ctx* c) {
  stage_output_plan(c);
  stage_hull(c);
  stage_finalize_state(c);
  stage_elements(c);
}

void redist_parts(gg_ctx* c) {
  for (int part = 0; part < 2; ++part) {
    SegTimer t(c, part == 0 ? kSegRedistXY : kSegRedistZ);
    over_classes(c, [&](const Launch& LL, int c0, int c1) {
      launch_redist(LL, part, c->sp, c->P2, din(c), dout(c), c->d_err, c->fp32, c0, c1);
    });
  }
}

// The redistancing over the output plan, then the input role takes the output
// boxes and their plan becomes the current one.
void stage_redistancing(gg_ctx* c) {
  redist_parts(c);
  { SegTimer t(c, kSegPost); launch_post_redist(launch_of(c), c->sp, c->d_state, din(c), dout(c)); }
  std::swap(c->P, c->P2);
  c->p2_valid = false;
}

// Level set and redistancing of one timestep with the hull on the side stream
// beside the redistancing: the hull reads the input role and IDLocal, the
// redistancing reads the input role and writes the output role, and the
// existence, volume and area finish after both.  The post-redistancing resize
// follows the deletion marks, as in the stage-by-stage order.
void stage_level_set_and_redistancing(gg_ctx* c) {
  if (c->timing) {
    stage_level_set(c);
    stage_redistancing(c);
    return;
  }
  Launch L = launch_of(c);
  stage_output_plan(c);
  if (c->topo.on) launch_topo_clear(L, c->sp, c->topo);
  hipEventRecord(c->ev_fork, c->stream);
  hipStreamWaitEvent(c->side, c->ev_fork, 0);
  Launch S = L;
  S.stream = c->side;
  launch_hull(S, c->sp, c->P, din(c), c->d_exflag, c->d_partials, c->IL[c->il_cur], c->topo, c->d_err, c->fp32);
  hipEventRecord(c->ev_join, c->side);
  for (int part = 0; part < 2; ++part) launch_redist(L, part, c->sp, c->P2, din(c), dout(c), c->d_err, c->fp32);
  hipStreamWaitEvent(c->stream, c->ev_join, 0);
  stage_finalize_state(c);
  stage_elements(c);
  launch_post_redist(L, c->sp, c->d_state, din(c), dout(c));
  std::swap(c->P, c->P2);
  c->p2_valid = false;
}

void stage_count(gg_ctx* c) {
  SegTimer t(c, kSegCount);
  launch_count(launch_of(c), c->sp, c->d_state, c->d_step, c->d_hist, c->hist_cap, c->d_err);
}

// one run_sim iteration body (src/grainhdl.cpp:836-919 without coarsening or output)
void enqueue_step(gg_ctx* c) {
  c->sel ^= 1;  // switchDistancebuffer in gridCoarsement's else branch
  if (c->timing) {
    stage_convolution(c);
    c->sel ^= 1;
    stage_lists(c);
  } else {
    // The list build reads the state, the box bounds and the current lists and
    // writes the other list buffer and the set-comparison start bounds; the
    // convolution reads and writes distance values and IDLocal bounds only. So
    // the lists go on the side stream, forked from and joined back to the main
    // stream (inside a captured graph too), beside the convolution.  In spatial
    // order the comparison's neighbour table (from the new lists) and its work
    // list (from the plan and the input bounds, which the convolution leaves
    // alone) follow them there.
    hipEventRecord(c->ev_fork, c->stream);
    hipStreamWaitEvent(c->side, c->ev_fork, 0);
    const Launch S{c->side, c->grid, c->grid / 8};
    launch_lists(S, c->sp, c->d_state, dout(c), din(c), c->LD, c->d_newb, c->d_err);  // the roles after the switch below
    if (c->cmp_order == GG_CMP_ORDER_SPACE) {
      launch_comparison_table(S, c->sp, c->d_state, dout(c), c->LD);
      launch_comparison_space(S, c->cmp_space_shape, c->sp, c->P, dout(c), din(c), c->LD, c->IL[c->il_cur],
                              c->IL[c->il_cur ^ 1], c->d_newb, c->W, c->d_err, c->fp32, kSpaceBuild);
      c->cmp_prebuilt = true;
    }
    hipEventRecord(c->ev_join, c->side);
    stage_convolution(c);
    c->sel ^= 1;
    hipStreamWaitEvent(c->stream, c->ev_join, 0);
  }
  stage_comparison(c);
  c->sel ^= 1;
  stage_level_set_and_redistancing(c);
  stage_count(c);
}

}  // namespace

extern "C" {

int gg_create(const gg_params* params, gg_ctx** out) {
  if (!params || !out) return GG_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GG_ERR_DEVICE;
  if (params->device < 0 || params->device >= ndev || params->device >= kScratchDevices) return GG_ERR_ARG;
  if (params->n_grains < 1 || params->n_grains > kMaxGrains || params->ngridpoints < 2 || params->grid_blowup < 0 ||
      (params->precision != GG_FP64_PLAIN && params->precision != GG_FP32_BATCHED))
    return GG_ERR_ARG;
  gg_ctx* c = new gg_ctx;
  c->p = *params;
  if (c->p.max_list <= 0) c->p.max_list = 512;
  if (c->p.max_list > 1024 || c->p.ngridpoints > kMaxGrid) {
    delete c;
    return GG_ERR_ARG;
  }
  c->fp32 = params->precision == GG_FP32_BATCHED;
  c->esz = c->fp32 ? 4 : 8;
  c->sp = SimParams{params->n_grains, params->ngridpoints, params->grid_blowup, params->boundary_grain_tube,
                    params->h, params->delta};
  if (hipSetDevice(params->device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return GG_ERR_HIP;
  }
  if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    gg_destroy(c);
    return GG_ERR_HIP;
  }
  {
    // this GPU's copy of the no-allocation scratch (see kNoAllocScratch)
    std::lock_guard<std::mutex> lk(g_scratch_mu[params->device]);
    if (!g_scratch_ptr[params->device]) {
      void* ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        gg_destroy(c);
        return GG_ERR_HIP;
      }
      g_scratch_ptr[params->device] = (int32_t*)ptr;
    }
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, params->device) == hipSuccess) c->grid = std::max(256, prop.multiProcessorCount * 8);
  const int N = params->n_grains;
  int rc;
  int32_t* dummy;
  (void)dummy;
  if ((rc = dalloc(c, &c->d_state, N + 1)) || (rc = dalloc(c, &c->d_newb, 6 * (N + 1))) ||
      (rc = dalloc(c, &c->d_exflag, N + 1)) || (rc = dalloc(c, &c->d_volume, N + 1)) ||
      (rc = dalloc(c, &c->d_area, N + 1)) || (rc = dalloc(c, &c->d_err, 1)) || (rc = dalloc(c, &c->d_step, 1)) ||
      (rc = dalloc(c, &c->d_hist, c->hist_cap)) || (rc = dalloc(c, &c->d_voxels, 1)) ||
      (rc = dalloc(c, &c->d_spec_off, N + 1))) {
    gg_destroy(c);
    return rc;
  }
  for (int r = 0; r < 2; ++r) {
    if ((rc = dalloc(c, &c->D[r].bounds, 6 * (N + 1))) || (rc = dalloc(c, &c->D[r].off, N + 1)) ||
        (rc = dalloc(c, &c->IL[r].bounds, 6 * (N + 1))) || (rc = dalloc(c, &c->IL[r].off, N + 1)) ||
        (rc = dalloc(c, &c->LD.buf[r], (size_t)(N + 1) * c->p.max_list)) || (rc = dalloc(c, &c->LD.len[r], N + 1))) {
      gg_destroy(c);
      return rc;
    }
  }
  c->LD.max_list = c->p.max_list;
  if ((rc = dalloc(c, &c->LD.tab, (size_t)(N + 1) * c->p.max_list)) ||
      (rc = dalloc(c, &c->LD.tab_id, (size_t)(N + 1) * c->p.max_list)) || (rc = dalloc(c, &c->LD.tab_len, N + 1))) {
    gg_destroy(c);
    return rc;
  }
  if ((rc = dalloc(c, &c->LD.sel, N + 1)) || (rc = dalloc(c, &c->P.cls_count, kClasses)) ||
      (rc = dalloc(c, &c->P.cls_start, kClasses)) || (rc = dalloc(c, &c->P.grains, N + 1)) ||
      (rc = dalloc(c, &c->P.grain_pos, N + 1)) || (rc = dalloc(c, &c->P.order, N)) || (rc = dalloc(c, &c->P.ncls, N)) ||
      (rc = dalloc(c, &c->P.totals, kNumQ * kClasses)) ||
      (rc = dalloc(c, &c->P.layer_base, kClasses)) || (rc = dalloc(c, &c->P.tile_base, kNumQ * kClasses)) ||
      (rc = dalloc(c, &c->d_partials, (size_t)4 * (N + 1) * kMaxN))) {
    gg_destroy(c);
    return rc;
  }
  for (int q = 0; q < kNumQ; ++q)
    if ((rc = dalloc(c, &c->P.pre[q], N + 1 + kClasses)) ||
        (rc = dalloc(c, &c->P.tmap[q], (size_t)kMaxTilesPerGrain * N))) {
      gg_destroy(c);
      return rc;
    }

  // the second plan shares the (constant) spatial grain order
  c->P2.order = c->P.order;
  if ((rc = dalloc(c, &c->P2.cls_count, kClasses)) || (rc = dalloc(c, &c->P2.cls_start, kClasses)) ||
      (rc = dalloc(c, &c->P2.grains, N + 1)) || (rc = dalloc(c, &c->P2.grain_pos, N + 1)) ||
      (rc = dalloc(c, &c->P2.ncls, N)) || (rc = dalloc(c, &c->P2.totals, kNumQ * kClasses)) ||
      (rc = dalloc(c, &c->P2.layer_base, kClasses)) || (rc = dalloc(c, &c->P2.tile_base, kNumQ * kClasses))) {
    gg_destroy(c);
    return rc;
  }
  for (int q = 0; q < kNumQ; ++q)
    if ((rc = dalloc(c, &c->P2.pre[q], N + 1 + kClasses)) ||
        (rc = dalloc(c, &c->P2.tmap[q], (size_t)kMaxTilesPerGrain * N))) {
      gg_destroy(c);
      return rc;
    }
  // zero the small per-grain state
  hipMemset(c->d_state, 0, sizeof(int32_t) * (N + 1));
  hipMemset(c->d_exflag, 0, sizeof(int32_t) * (N + 1));
  hipMemset(c->d_volume, 0, sizeof(double) * (N + 1));
  hipMemset(c->d_area, 0, sizeof(double) * (N + 1));
  hipMemset(c->d_err, 0, sizeof(int32_t));
  hipMemset(c->d_step, 0, sizeof(int32_t));
  hipMemset(c->LD.sel, 0, sizeof(int32_t) * (N + 1));
  hipMemset(c->LD.len[0], 0, sizeof(int32_t) * (N + 1));
  hipMemset(c->LD.len[1], 0, sizeof(int32_t) * (N + 1));
  if ((rc = build_tables(c))) {
    gg_destroy(c);
    return rc;
  }
  upload_mc_tables();
  if (hipDeviceSynchronize() != hipSuccess) {
    gg_destroy(c);
    return GG_ERR_HIP;
  }
  *out = c;
  return GG_OK;
}

int gg_destroy(gg_ctx* c) {
  if (!c) return GG_OK;
  hipSetDevice(c->p.device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->gexec) hipGraphExecDestroy(c->gexec);
  if (c->graph) hipGraphDestroy(c->graph);
  for (auto e : c->evpool) hipEventDestroy(e);
  for (void* p : c->allocs) hipFree(p);
  if (c->ev_fork) hipEventDestroy(c->ev_fork);
  if (c->ev_join) hipEventDestroy(c->ev_join);
  if (c->side) hipStreamDestroy(c->side);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return GG_OK;
}

const char* gg_last_error(const gg_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gg_set_boxes(gg_ctx* c, const int32_t* bounds, const uint8_t* exists) {
  if (!c || !bounds || !exists) return GG_ERR_ARG;
  hipSetDevice(c->p.device);
  const int N = c->p.n_grains, L = c->p.ngridpoints;
  std::vector<int32_t> b(6 * (N + 1), 0), st(N + 1, kDeleted);
  int64_t total = 0;
  for (int g = 1; g <= N; ++g) {
    if (!exists[g]) continue;
    int32_t* r = &b[6 * g];
    for (int a = 0; a < 6; ++a) r[a] = bounds[6 * g + a];
    if (r[3] <= r[0] || r[4] <= r[1] || r[5] <= r[2]) return fail(c, GG_ERR_ARG, "empty box bounds");
    cube_bounds_host(r, L);
    int n = r[3] - r[0];
    if (n > kMaxN) return fail(c, GG_ERR_CAPACITY, "box side above 96");
    st[g] = kAlive;
    total += (int64_t)n * n * n;
  }
  if (c->arena[0] == nullptr) {
    c->arena_cap = c->p.arena_voxels > 0 ? c->p.arena_voxels : 2 * total + 65536;
    c->spec_cap = c->arena_cap;
  } else if (total > c->arena_cap) {
    // the arenas, the spectral scratch and the work lists keep the size of the
    // first call
    return fail(c, GG_ERR_CAPACITY, "boxes exceed the arena sized by the first gg_set_boxes");
  }
  for (int r = 0; r < 2; ++r) {
    if (c->arena[r] == nullptr) {
      int rc;
      if (c->fp32) rc = dalloc(c, (float**)&c->arena[r], c->arena_cap);
      else rc = dalloc(c, (double**)&c->arena[r], c->arena_cap);
      if (rc) return rc;
      if ((rc = dalloc(c, &c->IL[r].arena, c->arena_cap))) return rc;
    }
    c->D[r].arena = c->arena[r];
  }
  if (!c->fp32 && c->d_spec == nullptr) {
    int rc = dalloc(c, (double2**)&c->d_spec, c->spec_cap);
    if (rc) return rc;
  }
  if (c->W.wl == nullptr) {
    // Wavefront tiles of a grain with side n, runs of >= 8 layers: at most
    // ceil(n^2/64) * ceil(n/8) <= n^3/128 for n >= 8, one below; the plan
    // keeps the sum of n^3 within the arena.
    c->W.bx = (L >> 4) / 8 + 1;  // cell coordinates of positions 0..L
    c->W.by = (L >> 3) / 8 + 1;
    c->W.nkeys = c->W.bx * c->W.by * c->W.by * 512;
    const int64_t cap = c->arena_cap / 128 + N + 64;
    if (cap > INT32_MAX) return fail(c, GG_ERR_CAPACITY, "comparison work list above 2^31 tiles");
    c->W.cap = (int)cap;
    int rc;
    if ((rc = dalloc(c, &c->W.wl, (size_t)cap)) || (rc = dalloc(c, &c->W.cnt, (size_t)c->W.nkeys + 1)) ||
        (rc = dalloc(c, &c->W.bsum, (size_t)c->W.nkeys / 4096 + 1)))
      return rc;
  }
  HIPCHK(c, hipMemcpy(c->d_state, st.data(), 4 * (N + 1), hipMemcpyHostToDevice));
  {
    // spatial order: Morton code of the cubed box centres (10 bits per axis)
    auto spread = [](uint64_t v) {
      v &= 0x3FF;
      v = (v | (v << 16)) & 0x30000FFull;
      v = (v | (v << 8)) & 0x300F00Full;
      v = (v | (v << 4)) & 0x30C30C3ull;
      v = (v | (v << 2)) & 0x9249249ull;
      return v;
    };
    std::vector<std::pair<uint64_t, int32_t>> key(N);
    for (int g = 1; g <= N; ++g) {
      const int32_t* r = &b[6 * g];
      uint64_t cx = (uint64_t)std::max(0, (r[0] + r[3]) / 2), cy = (uint64_t)std::max(0, (r[1] + r[4]) / 2),
               cz = (uint64_t)std::max(0, (r[2] + r[5]) / 2);
      uint64_t code = spread(cx) | (spread(cy) << 1) | (spread(cz) << 2);
      key[g - 1] = {st[g] == kAlive ? code : ~0ull, g};
    }
    std::sort(key.begin(), key.end());
    std::vector<int32_t> order(N);
    for (int i = 0; i < N; ++i) order[i] = key[i].second;
    HIPCHK(c, hipMemcpy(c->P.order, order.data(), 4 * N, hipMemcpyHostToDevice));
  }
  for (int r = 0; r < 2; ++r) {
    HIPCHK(c, hipMemcpy(c->D[r].bounds, b.data(), 4 * b.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->IL[r].bounds, b.data(), 4 * b.size(), hipMemcpyHostToDevice));
  }
  c->sel = 0;
  c->il_cur = 0;
  c->p2_valid = false;
  plan_layout(c, c->D[0]);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc = check_device_error(c);
  if (rc) return rc;
  for (int r = 0; r < 2; ++r) {
    if (r == 1) HIPCHK(c, hipMemcpy(c->D[1].off, c->D[0].off, 8 * (N + 1), hipMemcpyDeviceToDevice));
    HIPCHK(c, hipMemcpy(c->IL[r].off, c->D[0].off, 8 * (N + 1), hipMemcpyDeviceToDevice));
  }
  // IDLocal starts zero (std::vector value-initialisation, src/box.cpp:221)
  HIPCHK(c, hipMemset(c->IL[0].arena, 0, 4 * c->arena_cap));
  HIPCHK(c, hipMemset(c->IL[1].arena, 0, 4 * c->arena_cap));
  if (c->cd.ready) HIPCHK(c, hipMemset(c->cd.ready, 0, sizeof(int32_t)));  // no hull yet
  HIPCHK(c, hipDeviceSynchronize());
  c->boxes_set = true;
  return GG_OK;
}

int gg_init_distance(gg_ctx* c, const int32_t* idfield) {
  if (!c || !idfield) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  const int64_t L = c->p.ngridpoints;
  int32_t* d_id = nullptr;
  HIPCHK(c, hipMalloc(&d_id, sizeof(int32_t) * L * L * L));
  hipError_t e = hipMemcpy(d_id, idfield, sizeof(int32_t) * L * L * L, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(d_id);
    return fail(c, GG_ERR_HIP, std::string("hipMemcpy idfield: ") + hipGetErrorString(e));
  }
  launch_init_phi(launch_of(c), c->sp, c->P, din(c), d_id, c->fp32, 0, (int)L);
  // the redistancing runs over the output role's plan (same boxes here)
  plan_into(c, dout(c), c->P2);
  c->p2_valid = true;
  stage_redistancing(c);
  e = hipStreamSynchronize(c->stream);
  hipFree(d_id);
  if (e != hipSuccess) return fail(c, GG_ERR_HIP, std::string("init distance: ") + hipGetErrorString(e));
  return check_device_error(c);
}

int gg_init_distance_noalloc(gg_ctx* c, const int32_t* idfield) {
  if (!c || !idfield) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  int32_t* const scr = g_scratch_ptr[c->p.device];
  if (!scr) return fail(c, GG_ERR_STATE, "gg_init_distance_noalloc: no scratch on this device");
  const int64_t L = c->p.ngridpoints, layer = L * L;
  if (layer > kNoAllocScratch)
    return fail(c, GG_ERR_CAPACITY, "gg_init_distance_noalloc: ngridpoints^2 above the scratch's 4194304 entries");
  std::lock_guard<std::mutex> lk(g_scratch_mu[c->p.device]);
  const bool timing = c->timing;
  c->timing = false;  // stage timing would create events
  // slabs of whole grid layers: copy one into the scratch, write the initial
  // field of the box layers it holds; the next copy is ordered after that
  // kernel on the stream
  const int64_t per = kNoAllocScratch / layer;
  hipError_t e = hipSuccess;
  for (int64_t k0 = 0; k0 < L && e == hipSuccess; k0 += per) {
    const int64_t k1 = std::min(L, k0 + per);
    e = hipMemcpyAsync(scr, idfield + k0 * layer, sizeof(int32_t) * (size_t)((k1 - k0) * layer),
                       hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) launch_init_phi(launch_of(c), c->sp, c->P, din(c), scr, c->fp32, (int)k0, (int)k1);
  }
  if (e == hipSuccess) {
    plan_into(c, dout(c), c->P2);
    c->p2_valid = true;
    stage_redistancing(c);
    e = hipStreamSynchronize(c->stream);
  }
  c->timing = timing;
  if (e != hipSuccess) return fail(c, GG_ERR_HIP, std::string("init distance (no allocation): ") + hipGetErrorString(e));
  return check_device_error(c);
}

int gg_set_neighbour_lists(gg_ctx* c, const int32_t* offsets, const int32_t* ids) {
  if (!c || !offsets) return GG_ERR_ARG;
  hipSetDevice(c->p.device);
  const int N = c->p.n_grains, M = c->p.max_list;
  std::vector<int32_t> buf((size_t)(N + 1) * M, 0), len(N + 1, 0);
  for (int g = 1; g <= N; ++g) {
    int a = offsets[g], b = offsets[g + 1];
    if (b - a > M) return fail(c, GG_ERR_CAPACITY, "initial neighbour list longer than max_list");
    for (int i = a; i < b; ++i) {
      if (ids[i] < 1 || ids[i] > N || ids[i] == g) return fail(c, GG_ERR_ARG, "neighbour id out of range");
      buf[(size_t)g * M + (i - a)] = ids[i];
    }
    // the list builder relies on a list holding no id twice (as the reference's lists never do)
    std::vector<int32_t> sorted(ids + a, ids + b);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      return fail(c, GG_ERR_ARG, "initial neighbour list holds an id twice");
    len[g] = b - a;
  }
  std::vector<int32_t> zero(N + 1, 0);
  HIPCHK(c, hipMemcpy(c->LD.buf[0], buf.data(), 4 * buf.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->LD.len[0], len.data(), 4 * len.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->LD.sel, zero.data(), 4 * zero.size(), hipMemcpyHostToDevice));
  c->lists_set = true;
  return GG_OK;
}

int gg_build_neighbour_lists(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  if (!c->boxes_set) return fail(c, GG_ERR_STATE, "gg_set_boxes has not been called");
  hipSetDevice(c->p.device);
  launch_init_lists(launch_of(c), c->sp, c->d_state, din(c), c->LD, c->d_err);
  HIPCHK(c, hipGetLastError());
  c->lists_set = true;
  return gg_synchronize(c);
}

int gg_switch_buffers(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  c->sel ^= 1;
  return GG_OK;
}

int gg_convolution(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  stage_convolution(c);
  return GG_OK;
}

int gg_update_neighbour_lists(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  stage_lists(c);
  return GG_OK;
}

int gg_comparison(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  stage_comparison(c);
  return GG_OK;
}

// gridCoarsement's coarsening branch (src/grainhdl.cpp:1446-1492).  Enqueued
// on the context stream like the stages; ends with the switch that the next
// timestep's opening switch undoes.
int gg_coarsen_grid(gg_ctx* c, int32_t new_size, double new_dt) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  if (!c->lists_set) return fail(c, GG_ERR_STATE, "comparison lists have not been set");
  const int gb = c->p.grid_blowup;
  if (new_size < 1 || new_size + 2 * gb > c->p.ngridpoints || !(new_dt > 0))
    return fail(c, GG_ERR_ARG, "coarsening needs 1 <= new_size <= ngridpoints - 2*grid_blowup and dt > 0");
  hipSetDevice(c->p.device);
  if (!c->d_cnb && (rc = dalloc(c, &c->d_cnb, (size_t)6 * (c->p.n_grains + 1)))) return rc;
  // updateGridAndTimeVariables (src/grainhdl.cpp:1416-1444, Gaussian mode: dt from the caller)
  const double h_old = c->p.h, h_new = 1.0 / new_size;
  c->p.ngridpoints = new_size + 2 * gb;
  c->p.conv_domain_size = new_size;
  c->p.h = h_new;
  c->p.delta = gb / double(new_size);
  c->p.dt = new_dt;
  c->sp.ngp = c->p.ngridpoints;
  c->sp.h = c->p.h;
  c->sp.delta = c->p.delta;
  // the convolution tables depend on the grid and dt (the old ones are freed
  // once the work queued before this call is done); a captured step graph
  // holds the old grid in its kernel arguments
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamSynchronize(c->side));
  const TablesDev old_tb = c->TB;
  if ((rc = build_tables(c))) return rc;
  for (const void* q : {(const void*)old_tb.tw_off, (const void*)old_tb.twc, (const void*)old_tb.tws,
                        (const void*)old_tb.g_off, (const void*)old_tb.gtab, (const void*)old_tb.circ,
                        (const void*)old_tb.fplan})
    dfree(c, q);
  drop_graph(c);
  const CoarsenArgs a{gb, h_old / h_new, h_new / h_old, c->p.delta};
  Launch L = launch_of(c);
  // resizeGrid: the output buffer (the last redistancing's result) is
  // resampled into the input buffer at new bounds and a new layout; then the
  // output buffer takes the input buffer's bounds
  launch_coarsen_bounds(L, c->sp, a, c->d_state, dout(c).bounds, din(c).bounds, c->d_cnb);
  plan_layout(c, din(c));
  {
    SegTimer t(c, kSegResample);
    launch_coarsen_resample(L, c->sp, a, c->d_state, dout(c), din(c), c->fp32);
  }
  launch_post_redist(L, c->sp, c->d_state, dout(c), din(c));
  // recalculateIDLocal (src/box.cpp:2699-2702): IDLocal to the output bounds,
  // then the comparison over the lists of the last comparison
  launch_idl_resize(L, c->sp, c->d_state, dout(c), c->IL[c->il_cur ^ 1]);
  launch_comparison_table(L, c->sp, c->d_state, din(c), c->LD, true);
  for (int cls = 0; cls < kClasses; ++cls)
    launch_comparison(L, cls, c->cmp_variant[cls], c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur], c->IL[c->il_cur ^ 1],
                      c->d_newb, c->fp32);
  c->il_cur ^= 1;
  // extractContour on the resampled input buffer
  if (c->topo.on) launch_topo_clear(L, c->sp, c->topo);
  launch_hull(L, c->sp, c->P, din(c), c->d_exflag, c->d_partials, c->IL[c->il_cur], c->topo, c->d_err, c->fp32);
  launch_finalize(L, c->sp, c->P, c->d_state, c->LD, c->d_exflag, c->d_partials, c->d_cnb, din(c), dout(c),
                  c->d_volume, c->d_area, kFinBounds | kFinState | kFinCoarsen);
  stage_elements(c);
  plan_layout(c, dout(c));
  c->p2_valid = false;
  c->sel ^= 1;
  return GG_OK;
}

int gg_level_set(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  stage_level_set(c);
  return GG_OK;
}

int gg_redistancing(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  // The redistancing writes the output boxes the level set's first part laid
  // out (their plan); a second call, or one after a coarsening, has none.
  if (!c->p2_valid)
    return fail(c, GG_ERR_STATE, "gg_redistancing needs the output boxes of a gg_level_set first");
  hipSetDevice(c->p.device);
  stage_redistancing(c);
  return GG_OK;
}

int gg_count_grains(gg_ctx* c, int32_t* count) {
  if (!c || !count) return GG_ERR_ARG;
  hipSetDevice(c->p.device);
  stage_count(c);
  int rc = gg_synchronize(c);
  if (rc) return rc;
  int32_t st = 0;
  HIPCHK(c, hipMemcpy(&st, c->d_step, 4, hipMemcpyDeviceToHost));
  if (st < 1 || st > c->hist_cap) return fail(c, GG_ERR_CAPACITY, "count history full");
  HIPCHK(c, hipMemcpy(count, c->d_hist + st - 1, 4, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_timestep(gg_ctx* c, int32_t nsteps, int32_t use_graph) {
  if (!c || nsteps < 0) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  if (!c->lists_set) return fail(c, GG_ERR_STATE, "gg_set_neighbour_lists has not been called");
  hipSetDevice(c->p.device);
  if (!use_graph || c->timing) {
    for (int s = 0; s < nsteps; ++s) enqueue_step(c);
    return GG_OK;
  }
  int s = 0;
  // the graph holds two timesteps (an even number of role swaps) and is valid
  // for the role parity it was captured at
  if (c->gexec && (c->graph_sel != c->sel || c->graph_il != c->il_cur || c->graph_plan != c->P.grains)) {
    if (nsteps > 0) {
      enqueue_step(c);
      ++s;
    }
  }
  if (!c->gexec && nsteps - s >= 2) {
    c->graph_sel = c->sel;
    c->graph_il = c->il_cur;
    c->graph_plan = c->P.grains;
    const bool p2v = c->p2_valid;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(c);
    enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
    if (e != hipSuccess) return fail(c, GG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail(c, GG_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    // capturing advanced the host-side role state by two steps; restore it
    c->sel = c->graph_sel;
    c->il_cur = c->graph_il;
    if (c->P.grains != c->graph_plan) std::swap(c->P, c->P2);
    c->p2_valid = p2v;
  }
  while (nsteps - s >= 2 && c->gexec && c->graph_sel == c->sel && c->graph_il == c->il_cur &&
         c->graph_plan == c->P.grains) {
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    s += 2;
  }
  for (; s < nsteps; ++s) enqueue_step(c);
  return GG_OK;
}

int gg_autotune(gg_ctx* c, int32_t reps, int32_t* choice, float* ms) {
  if (!c || reps < 1) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  if (!c->lists_set) return fail(c, GG_ERR_STATE, "gg_set_neighbour_lists has not been called");
  hipSetDevice(c->p.device);
  if ((rc = gg_synchronize(c))) return rc;
  const bool timing = c->timing;
  c->timing = false;
  // one timestep up to the comparison's inputs
  c->sel ^= 1;
  stage_convolution(c);
  c->sel ^= 1;
  stage_lists(c);
  launch_comparison_table(launch_of(c), c->sp, c->d_state, din(c), c->LD);
  // The comparison is a function of its inputs (the input role, the previous
  // IDLocal, the lists; the bounds atomics are minima and maxima), so every
  // variant can run on the same inputs: time each, per size class.
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  const bool space = c->cmp_order == GG_CMP_ORDER_SPACE;
  for (int cls = 0; cls < (space ? 1 : kClasses); ++cls) {
    float best = 0.f;
    int bv = space ? c->cmp_space_shape : c->cmp_variant[cls];
    // spatial order: the shapes of its one launch (work list build included)
    auto run = [&](int v) {
      if (space)
        launch_comparison_space(launch_of(c), v, c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur],
                                c->IL[c->il_cur ^ 1], c->d_newb, c->W, c->d_err, c->fp32);
      else
        launch_comparison(launch_of(c), cls, v, c->sp, c->P, din(c), dout(c), c->LD, c->IL[c->il_cur],
                          c->IL[c->il_cur ^ 1], c->d_newb, c->fp32);
    };
    for (int v = 0; v < (space ? kCmpSpaceShapes : kCmpVariants); ++v) {
      run(v);  // warm
      hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps; ++r) run(v);
      hipEventRecord(e1, c->stream);
      hipError_t e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      if (e != hipSuccess) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        c->timing = timing;
        return fail(c, GG_ERR_HIP, std::string("autotune: ") + hipGetErrorString(e));
      }
      t /= (float)reps;
      if (ms) ms[cls * kCmpVariants + v] = t;
      if (v == 0 || t < best) {
        best = t;
        bv = v;
      }
    }
    if (space) {
      c->cmp_space_shape = bv;
      if (choice)
        for (int k = 0; k < kClasses; ++k) choice[k] = bv;
    } else {
      c->cmp_variant[cls] = bv;
      if (choice) choice[cls] = bv;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  // the rest of the timestep, as gg_timestep runs it
  c->il_cur ^= 1;
  c->sel ^= 1;
  stage_level_set_and_redistancing(c);
  stage_count(c);
  c->timing = timing;
  drop_graph(c);  // a captured step holds the old shapes
  return gg_synchronize(c);
}

int gg_set_comparison_order(gg_ctx* c, int32_t order, int32_t shape) {
  if (!c) return GG_ERR_ARG;
  if (order != GG_CMP_ORDER_GRAIN && order != GG_CMP_ORDER_SPACE)
    return fail(c, GG_ERR_ARG, "unknown comparison order");
  if (shape < 0 || shape >= kCmpSpaceShapes) return fail(c, GG_ERR_ARG, "spatial tile shape out of range");
  c->cmp_order = order;
  c->cmp_space_shape = shape;
  drop_graph(c);
  return GG_OK;
}

int gg_get_comparison_order(gg_ctx* c, int32_t* order, int32_t* shape) {
  if (!c || !order) return GG_ERR_ARG;
  *order = c->cmp_order;
  if (shape) *shape = c->cmp_space_shape;
  return GG_OK;
}

int gg_get_tuning(gg_ctx* c, int32_t* variants) {
  if (!c || !variants) return GG_ERR_ARG;
  for (int k = 0; k < kClasses; ++k) variants[k] = c->cmp_variant[k];
  return GG_OK;
}

int gg_set_convolution_form(gg_ctx* c, int32_t form) {
  if (!c) return GG_ERR_ARG;
  if (form != GG_CONV_DEFAULT && form != GG_CONV_SPLIT_BF16) return fail(c, GG_ERR_ARG, "unknown convolution form");
  if (form == GG_CONV_SPLIT_BF16 && !c->fp32)
    return fail(c, GG_ERR_ARG, "the split-BF16 convolution is a form of the FP32 batched path");
  c->conv_bf16 = form == GG_CONV_SPLIT_BF16;
  drop_graph(c);
  return GG_OK;
}

int gg_set_work_order(gg_ctx* c, int32_t order) {
  if (!c) return GG_ERR_ARG;
  if (order != GG_ORDER_SPATIAL && order != GG_ORDER_VOLUME) return fail(c, GG_ERR_ARG, "unknown work order");
  int32_t* v = nullptr;
  if (order == GG_ORDER_VOLUME) {
    if (!c->d_vorder) {
      int rc = dalloc(c, &c->d_vorder, (size_t)c->p.n_grains);
      if (rc) return rc;
    }
    v = c->d_vorder;
  }
  c->P.vorder = v;
  c->P2.vorder = v;
  drop_graph(c);
  return GG_OK;
}

int gg_get_work_order(gg_ctx* c, int32_t* order) {
  if (!c || !order) return GG_ERR_ARG;
  *order = c->P.vorder ? GG_ORDER_VOLUME : GG_ORDER_SPATIAL;
  return GG_OK;
}

int gg_set_tuning(gg_ctx* c, const int32_t* variants) {
  if (!c || !variants) return GG_ERR_ARG;
  for (int k = 0; k < kClasses; ++k)
    if (variants[k] < 0 || variants[k] >= kCmpVariants) return fail(c, GG_ERR_ARG, "tuning variant out of range");
  for (int k = 0; k < kClasses; ++k) c->cmp_variant[k] = variants[k];
  drop_graph(c);
  return GG_OK;
}

int gg_synchronize(gg_ctx* c) {
  if (!c) return GG_ERR_ARG;
  hipSetDevice(c->p.device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());  // a kernel launch that failed (e.g. resources) since the last check
  if (c->timing) collect_timing(c);
  if (!c->boxes_set) return GG_OK;
  return check_device_error(c);
}

int gg_get_count_history(gg_ctx* c, int32_t* counts, int32_t cap, int32_t* n) {
  if (!c || !n) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  int32_t st = 0;
  HIPCHK(c, hipMemcpy(&st, c->d_step, 4, hipMemcpyDeviceToHost));
  *n = st;
  int m = std::min(std::min(st, cap), c->hist_cap);
  if (counts && m > 0) HIPCHK(c, hipMemcpy(counts, c->d_hist, 4 * m, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_get_grain_stats(gg_ctx* c, int32_t* state, double* volume, double* area) {
  if (!c) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const int N = c->p.n_grains;
  if (state) HIPCHK(c, hipMemcpy(state, c->d_state, 4 * (N + 1), hipMemcpyDeviceToHost));
  if (volume) HIPCHK(c, hipMemcpy(volume, c->d_volume, 8 * (N + 1), hipMemcpyDeviceToHost));
  if (area) HIPCHK(c, hipMemcpy(area, c->d_area, 8 * (N + 1), hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_get_bounds(gg_ctx* c, int32_t id, int32_t which, int32_t* b6) {
  if (!c || !b6 || id < 0 || id > c->p.n_grains) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const DescDev& d = which ? dout(c) : din(c);
  HIPCHK(c, hipMemcpy(b6, d.bounds + 6 * id, 24, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_download_distance(gg_ctx* c, int32_t id, int32_t which, double* dst, int64_t cap, int64_t* n) {
  if (!c || !n || id < 1 || id > c->p.n_grains) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const DescDev& d = which ? dout(c) : din(c);
  int32_t b[6];
  int64_t off = 0;
  HIPCHK(c, hipMemcpy(b, d.bounds + 6 * id, 24, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&off, d.off + id, 8, hipMemcpyDeviceToHost));
  const int64_t cnt = (int64_t)(b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
  *n = cnt;
  if (!dst) return GG_OK;
  if (cnt > cap) return fail(c, GG_ERR_ARG, "destination too small");
  if (c->fp32) {
    std::vector<float> tmp(cnt);
    HIPCHK(c, hipMemcpy(tmp.data(), (float*)d.arena + off, 4 * cnt, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < cnt; ++i) dst[i] = tmp[i];
  } else {
    HIPCHK(c, hipMemcpy(dst, (double*)d.arena + off, 8 * cnt, hipMemcpyDeviceToHost));
  }
  return GG_OK;
}

int gg_upload_phi(gg_ctx* c, int32_t id, int32_t which, const double* src, int64_t n) {
  if (!c || !src || id < 1 || id > c->p.n_grains || (which != 0 && which != 1)) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  if ((rc = gg_synchronize(c))) return rc;  // the stream's work on the buffer is done before it is overwritten
  int32_t st = 0;
  HIPCHK(c, hipMemcpy(&st, c->d_state + id, 4, hipMemcpyDeviceToHost));
  if (st != kAlive) return fail(c, GG_ERR_STATE, "gg_upload_phi: grain is not alive");
  const DescDev& d = which ? dout(c) : din(c);
  int32_t b[6];
  int64_t off = 0;
  HIPCHK(c, hipMemcpy(b, d.bounds + 6 * id, 24, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&off, d.off + id, 8, hipMemcpyDeviceToHost));
  const int64_t cnt = (int64_t)(b[3] - b[0]) * (b[4] - b[1]) * (b[5] - b[2]);
  if (n != cnt) return fail(c, GG_ERR_ARG, "gg_upload_phi: element count differs from the buffer's box volume");
  if (c->fp32) {
    std::vector<float> tmp(cnt);
    for (int64_t i = 0; i < cnt; ++i) tmp[i] = (float)src[i];
    HIPCHK(c, hipMemcpy((float*)d.arena + off, tmp.data(), 4 * cnt, hipMemcpyHostToDevice));
  } else {
    HIPCHK(c, hipMemcpy((double*)d.arena + off, src, 8 * cnt, hipMemcpyHostToDevice));
  }
  return GG_OK;
}

int gg_download_idlocal(gg_ctx* c, int32_t id, int32_t* b6, uint32_t* dst, int64_t cap, int64_t* n) {
  if (!c || !n || !b6 || id < 1 || id > c->p.n_grains) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const IdlDev& il = c->IL[c->il_cur];
  int64_t off = 0;
  HIPCHK(c, hipMemcpy(b6, il.bounds + 6 * id, 24, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(&off, il.off + id, 8, hipMemcpyDeviceToHost));
  const int64_t cnt = (int64_t)(b6[3] - b6[0]) * (b6[4] - b6[1]) * (b6[5] - b6[2]);
  *n = cnt;
  if (!dst) return GG_OK;
  if (cnt > cap) return fail(c, GG_ERR_ARG, "destination too small");
  HIPCHK(c, hipMemcpy(dst, il.arena + off, 4 * cnt, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_download_list(gg_ctx* c, int32_t id, int32_t which, int32_t* dst, int32_t cap, int32_t* n) {
  if (!c || !n || id < 1 || id > c->p.n_grains) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  int32_t sel = 0, len = 0;
  HIPCHK(c, hipMemcpy(&sel, c->LD.sel + id, 4, hipMemcpyDeviceToHost));
  const int b = which ? sel ^ 1 : sel;
  HIPCHK(c, hipMemcpy(&len, c->LD.len[b] + id, 4, hipMemcpyDeviceToHost));
  *n = len;
  if (!dst) return GG_OK;
  if (len > cap) return fail(c, GG_ERR_ARG, "destination too small");
  if (len > 0)
    HIPCHK(c, hipMemcpy(dst, c->LD.buf[b] + (size_t)id * c->p.max_list, 4 * len, hipMemcpyDeviceToHost));
  return GG_OK;
}

static int alloc_topology(gg_ctx* c) {
  if (c->topo.key) return GG_OK;
  const size_t n = (size_t)(c->p.n_grains + 1) * kTopoCap;
  int rc;
  const size_t nj = (size_t)(c->p.n_grains + 1) * kJuncCap;
  if ((rc = dalloc(c, &c->topo.key, n)) || (rc = dalloc(c, &c->topo.ord, n)) || (rc = dalloc(c, &c->topo.jkey, nj)) ||
      (rc = dalloc(c, &c->topo.jpos, 3 * nj)) || (rc = dalloc(c, &c->topo.jcnt, nj)) ||
      (rc = dalloc(c, &c->topo.jarea, nj)))
    return rc;
  // empty tables: no grain touched yet
  HIPCHK(c, hipMemset(c->topo.key, 0xFF, n * sizeof(int32_t)));
  HIPCHK(c, hipMemset(c->topo.jkey, 0, nj * sizeof(unsigned long long)));
  HIPCHK(c, hipMemset(c->topo.jcnt, 0, nj * sizeof(int32_t)));
  HIPCHK(c, hipMemset(c->topo.jpos, 0, 3 * nj * sizeof(unsigned long long)));
  HIPCHK(c, hipMemset(c->topo.jarea, 0, nj * sizeof(double)));
  return GG_OK;
}

int gg_enable_hull_topology(gg_ctx* c, int32_t enable) {
  if (!c) return GG_ERR_ARG;
  hipSetDevice(c->p.device);
  c->topo_user = enable ? 1 : 0;
  const int on = (c->topo_user || c->corr_on) ? 1 : 0;
  if (on) {
    int rc = alloc_topology(c);
    if (rc) return rc;
  }
  if (on != c->topo.on) {
    c->topo.on = on;
    drop_graph(c);  // a captured step graph holds the old setting
  }
  return GG_OK;
}

int gg_set_physics(gg_ctx* c, const gg_physics* ph, const double* quat, const double* see) {
  if (!c || !ph) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipStreamSynchronize(c->side));
  drop_graph(c);  // kernel arguments change
  const int N = c->p.n_grains;
  if (!ph->correction) {
    c->corr_on = false;
    c->topo.corr = 0;
    c->topo.on = c->topo_user;
    return GG_OK;
  }
  const bool aniso = ph->mobility_model || !ph->unique_energies;
  if (aniso && !quat) return fail(c, GG_ERR_ARG, "orientations are needed for misorientation-dependent energies or mobilities");
  if (ph->stored_energy && !see) return fail(c, GG_ERR_ARG, "stored energies are needed with stored_energy on");
  if (aniso && N > GG_MAX_ANISO_GRAINS)
    return fail(c, GG_ERR_CAPACITY, "misorientation-dependent energies need n_grains <= GG_MAX_ANISO_GRAINS");
  if ((rc = alloc_topology(c))) return rc;
  CorrDev& d = c->cd;
  const size_t n1 = (size_t)N + 1;
  if (!d.ready) {
    const size_t ne = n1 * kElemCap;
    const int64_t cap = c->arena_cap / 8 + 65536;  // triangles of one hull pass (about 3 per surface voxel)
    const int64_t tcap = (cap + cap / 4) / kChunk * kChunk;  // the same, grouped, with each element's padding
    if ((rc = dalloc(c, &c->topo.jord, n1 * kJuncCap)) || (rc = dalloc(c, &c->topo.raw, (size_t)cap)) ||
        (rc = dalloc(c, &c->topo.raw_n, 1)) || (rc = dalloc(c, &d.ready, 1)) || (rc = dalloc(c, &d.ecount, n1)) ||
        (rc = dalloc(c, &d.eclass, 4 * n1)) || (rc = dalloc(c, &d.epos, 3 * ne)) || (rc = dalloc(c, &d.emask, ne)) ||
        (rc = dalloc(c, &d.energy, ne)) || (rc = dalloc(c, &d.mobility, ne)) ||
        (rc = dalloc(c, &d.efirst, n1 * (kElemCap + 1))) || (rc = dalloc(c, &d.ecur, ne)) ||
        (rc = dalloc(c, &d.slot_elem, n1 * kJuncCap)) || (rc = dalloc(c, &d.tbase, n1)) ||
        (rc = dalloc(c, &d.tri, 3 * (size_t)tcap)) || (rc = dalloc(c, &d.cbox, 6 * (size_t)(tcap / kChunk))) ||
        (rc = dalloc(c, &d.tri_n, 1)) || (rc = dalloc(c, (double**)&d.see, n1)))
      return rc;
    c->topo.raw_cap = cap;
    d.tri_cap = tcap;
    HIPCHK(c, hipMemset(d.ecount, 0, 4 * n1));
    HIPCHK(c, hipMemset(d.slot_elem, 0xFF, 4 * n1 * kJuncCap));
  }
  if (aniso) {
    if (!d.gam) {
      if ((rc = dalloc(c, (double**)&d.gam, n1 * n1)) || (rc = dalloc(c, (double**)&d.mob, n1 * n1))) return rc;
    }
    std::vector<double> gam(n1 * n1), mob(n1 * n1);
    gg_gb_pair_tables((int32_t)n1, quat, ph->mobility_model, ph->unique_energies, gam.data(), mob.data(), 0);
    HIPCHK(c, hipMemcpy((void*)d.gam, gam.data(), 8 * n1 * n1, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy((void*)d.mob, mob.data(), 8 * n1 * n1, hipMemcpyHostToDevice));
  }
  std::vector<double> sv(n1, 0.0);
  if (ph->stored_energy)
    for (size_t g = 0; g < n1; ++g) sv[g] = see[g];
  HIPCHK(c, hipMemcpy((void*)d.see, sv.data(), 8 * n1, hipMemcpyHostToDevice));
  // the correction starts after the next hull (run_sim's loop != 0)
  HIPCHK(c, hipMemset(d.ready, 0, sizeof(int32_t)));
  d.see_on = ph->stored_energy ? 1 : 0;
  d.aniso = aniso ? 1 : 0;
  const double time_slope = 1 / 1.21;  // grainhdl::TimeSlope in the Gaussian convolution mode
  d.tsq = time_slope * time_slope;
  c->tl_drag = ph->triple_line_drag;
  c->corr_on = true;
  c->topo.corr = 1;
  c->topo.on = 1;
  return GG_OK;
}

int gg_download_elements(gg_ctx* c, int32_t id, int32_t* cls, int32_t* nnb, int32_t* nb8, int32_t* nbary,
                         double* pos, double* energy, double* mobility, int32_t cap, int32_t* n) {
  if (!c || !n || id < 1 || id > c->p.n_grains) return GG_ERR_ARG;
  if (!c->cd.ready) return fail(c, GG_ERR_STATE, "the convolution correction was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  int32_t E = 0, ec[4] = {0, 0, 0, 0};
  HIPCHK(c, hipMemcpy(&E, c->cd.ecount + id, 4, hipMemcpyDeviceToHost));
  *n = E;
  if (!cls) return GG_OK;
  if (E > cap || !nnb || !nb8 || !nbary || !pos || !energy || !mobility)
    return fail(c, GG_ERR_ARG, "destination too small");
  if (E == 0) return GG_OK;
  const int64_t eb = (int64_t)id * kElemCap;
  std::vector<int32_t> key(kTopoCap), cnt(E);
  std::vector<unsigned long long> em(E);
  HIPCHK(c, hipMemcpy(ec, c->cd.eclass + 4 * id, 12, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(key.data(), c->topo.key + (int64_t)id * kTopoCap, 4 * kTopoCap, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(em.data(), c->cd.emask + eb, 8 * E, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cnt.data(), c->cd.ecur + eb, 4 * E, hipMemcpyDeviceToHost));  // barycenters scattered
  HIPCHK(c, hipMemcpy(pos, c->cd.epos + 3 * eb, 24 * E, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(energy, c->cd.energy + eb, 8 * E, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(mobility, c->cd.mobility + eb, 8 * E, hipMemcpyDeviceToHost));
  for (int e = 0; e < E; ++e) {
    cls[e] = e < ec[0] ? 3 : e < ec[0] + ec[1] ? 2 : e < ec[0] + ec[1] + ec[2] ? 1 : 0;
    std::vector<int64_t> ids;
    for (int b = 0; b < kTopoCap; ++b)
      if ((em[e] >> b) & 1ull) ids.push_back((int64_t)(uint32_t)key[b]);
    std::sort(ids.begin(), ids.end());
    nnb[e] = (int32_t)ids.size();
    for (int k = 0; k < 8; ++k) nb8[8 * e + k] = k < (int)ids.size() ? (int32_t)ids[k] : -1;
    nbary[e] = cnt[e];
  }
  return GG_OK;
}

// The grain's topology table, sorted by first appearance.
static int topo_ids(gg_ctx* c, int32_t id, std::vector<std::pair<int32_t, int32_t>>& out) {
  out.clear();
  std::vector<int32_t> k(kTopoCap), o(kTopoCap);
  HIPCHK(c, hipMemcpy(k.data(), c->topo.key + (size_t)id * kTopoCap, kTopoCap * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(o.data(), c->topo.ord + (size_t)id * kTopoCap, kTopoCap * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int i = 0; i < kTopoCap; ++i)
    if (k[i] != -1) out.push_back({o[i], k[i]});
  std::sort(out.begin(), out.end());
  return GG_OK;
}

int gg_get_hull_neighbour_counts(gg_ctx* c, int32_t* counts, uint8_t* touches_wall) {
  if (!c || !counts || !touches_wall) return GG_ERR_ARG;
  if (!c->topo.key) return fail(c, GG_ERR_STATE, "hull topology was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const size_t n = (size_t)(c->p.n_grains + 1) * kTopoCap;
  std::vector<int32_t> k(n);
  HIPCHK(c, hipMemcpy(k.data(), c->topo.key, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int g = 0; g <= c->p.n_grains; ++g) {
    int cnt = 0, wall = 0;
    for (int i = 0; i < kTopoCap; ++i) {
      const int32_t v = k[(size_t)g * kTopoCap + i];
      cnt += v != -1;
      wall |= v == 0;
    }
    counts[g] = cnt;
    touches_wall[g] = (uint8_t)wall;
  }
  return GG_OK;
}

int gg_get_hull_junction_counts(gg_ctx* c, int32_t* counts4) {
  if (!c || !counts4) return GG_ERR_ARG;
  if (!c->topo.key) return fail(c, GG_ERR_STATE, "hull topology was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const size_t nj = (size_t)(c->p.n_grains + 1) * kJuncCap;
  std::vector<unsigned long long> j(nj);
  HIPCHK(c, hipMemcpy(j.data(), c->topo.jkey, nj * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  for (int g = 0; g <= c->p.n_grains; ++g) {
    int32_t* o = counts4 + 4 * (size_t)g;
    o[0] = o[1] = o[2] = o[3] = 0;
    for (int i = 0; i < kJuncCap; ++i) {
      const unsigned long long m = j[(size_t)g * kJuncCap + i];
      if (!m) continue;
      const int k = __builtin_popcountll(m);
      ++o[k == 2 ? 0 : k == 3 ? 1 : k == 4 ? 2 : 3];
    }
  }
  return GG_OK;
}

int gg_get_element_counts(gg_ctx* c, int32_t* counts4) {
  if (!c || !counts4) return GG_ERR_ARG;
  if (!c->cd.ready) return fail(c, GG_ERR_STATE, "the convolution correction was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const size_t n1 = (size_t)c->p.n_grains + 1;
  std::vector<int32_t> e(n1), cls(4 * n1);
  HIPCHK(c, hipMemcpy(e.data(), c->cd.ecount, 4 * n1, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cls.data(), c->cd.eclass, 16 * n1, hipMemcpyDeviceToHost));
  for (size_t g = 0; g < n1; ++g) {
    const int32_t E = e[g];
    const int32_t nh = E ? cls[4 * g] : 0, nq = E ? cls[4 * g + 1] : 0, nt = E ? cls[4 * g + 2] : 0;
    counts4[4 * g] = E - nh - nq - nt;  // grain boundaries
    counts4[4 * g + 1] = nt;
    counts4[4 * g + 2] = nq;
    counts4[4 * g + 3] = nh;
  }
  return GG_OK;
}

int gg_get_faces(gg_ctx* c, int32_t* grain_a, int32_t* grain_b, double* area, int64_t cap, int64_t* n) {
  if (!c || !n) return GG_ERR_ARG;
  if (!c->topo.key) return fail(c, GG_ERR_STATE, "hull topology was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const int N = c->p.n_grains;
  const size_t nk = (size_t)(N + 1) * kTopoCap, nj = (size_t)(N + 1) * kJuncCap;
  std::vector<int32_t> key(nk), st(N + 1);
  std::vector<unsigned long long> jk(nj);
  std::vector<double> ja(nj);
  HIPCHK(c, hipMemcpy(key.data(), c->topo.key, nk * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(jk.data(), c->topo.jkey, nj * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(ja.data(), c->topo.jarea, nj * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(st.data(), c->d_state, (N + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  const double h2 = c->p.h * c->p.h;
  int64_t cnt = 0;
  std::vector<std::pair<int32_t, double>> mine;
  for (int g = 1; g <= N; ++g) {
    if (st[g] != kAlive) continue;
    const int32_t* kg = &key[(size_t)g * kTopoCap];
    mine.clear();
    for (int i = 0; i < kJuncCap; ++i) {
      const unsigned long long m = jk[(size_t)g * kJuncCap + i];
      if (__builtin_popcountll(m) != 2) continue;  // a grain boundary: the grain and one neighbour
      const int b0 = __builtin_ctzll(m), b1 = 63 - __builtin_clzll(m);
      const int32_t a0 = kg[b0], a1 = kg[b1];
      const int32_t nb = a0 == g ? a1 : a0;
      if ((a0 != g && a1 != g) || g <= nb) continue;  // get_Faces: only owner id > neighbour id
      mine.push_back({nb, ja[(size_t)g * kJuncCap + i] * h2});
    }
    std::sort(mine.begin(), mine.end());
    for (const auto& f : mine) {
      if (grain_a && cnt < cap) {
        grain_a[cnt] = g;
        grain_b[cnt] = f.first;
        area[cnt] = f.second;
      }
      ++cnt;
    }
  }
  *n = cnt;
  return grain_a && cnt > cap ? fail(c, GG_ERR_CAPACITY, "faces: cap too small") : GG_OK;
}

int gg_download_hull_neighbours(gg_ctx* c, int32_t id, int32_t* dst, int32_t cap, int32_t* n) {
  if (!c || !n || id < 0 || id > c->p.n_grains) return GG_ERR_ARG;
  if (!c->topo.key) return fail(c, GG_ERR_STATE, "hull topology was never enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  std::vector<std::pair<int32_t, int32_t>> v;
  if ((rc = topo_ids(c, id, v))) return rc;
  *n = (int32_t)v.size();
  if (dst)
    for (int i = 0; i < (int)v.size() && i < cap; ++i) dst[i] = v[i].second;
  return (int)v.size() > cap && dst ? fail(c, GG_ERR_CAPACITY, "hull neighbours: cap too small") : GG_OK;
}

int gg_voxel_container(gg_ctx* c, uint32_t* dst, int64_t cap, int32_t* size) {
  if (!c || !size) return GG_ERR_ARG;
  int rc = require_boxes(c);
  if (rc) return rc;
  hipSetDevice(c->p.device);
  const int L = c->p.conv_domain_size;
  *size = L;
  const int64_t n = (int64_t)L * L * L;
  if (!dst) return GG_OK;
  if (n > cap) return fail(c, GG_ERR_ARG, "destination too small");
  if (c->cont_cap < n) {
    // a context that once wrote a container keeps the buffer (a later
    // coarsening only shrinks the grid); every call ends synchronised, so a
    // smaller old buffer is idle here
    if (c->d_cont) dfree(c, c->d_cont);
    c->d_cont = nullptr;
    c->cont_cap = 0;
    if ((rc = dalloc(c, &c->d_cont, (size_t)n))) return rc;
    c->cont_cap = n;
  }
  {
    SegTimer t(c, kSegContainer);
    launch_container(launch_of(c), c->sp, c->d_state, dout(c), c->d_cont, L, c->fp32);
  }
  if ((rc = gg_synchronize(c))) return rc;
  HIPCHK(c, hipMemcpy(dst, c->d_cont, 4 * (size_t)n, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_download_hull_elements(gg_ctx* c, int32_t id, int32_t* nids, int32_t* ids8, int32_t* ntri, double* pos,
                              int32_t cap, int32_t* n) {
  if (!c || !n || id < 1 || id > c->p.n_grains) return GG_ERR_ARG;
  if (!c->topo.key) return fail(c, GG_ERR_STATE, "hull topology is not enabled");
  int rc = gg_synchronize(c);
  if (rc) return rc;
  std::vector<int32_t> key(kTopoCap), cnt(kJuncCap);
  std::vector<unsigned long long> jk(kJuncCap);
  std::vector<unsigned long long> jp(3 * kJuncCap);
  HIPCHK(c, hipMemcpy(key.data(), c->topo.key + (int64_t)id * kTopoCap, 4 * kTopoCap, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(jk.data(), c->topo.jkey + (int64_t)id * kJuncCap, 8 * kJuncCap, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(cnt.data(), c->topo.jcnt + (int64_t)id * kJuncCap, 4 * kJuncCap, hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(jp.data(), c->topo.jpos + (int64_t)id * kJuncCap * 3, 24 * kJuncCap, hipMemcpyDeviceToHost));
  struct Rec {
    std::vector<int64_t> ids;
    int32_t nt;
    double p[3];
  };
  std::vector<Rec> recs;
  for (int s = 0; s < kJuncCap; ++s) {
    if (!jk[s] || cnt[s] <= 0) continue;
    Rec r;
    for (int b = 0; b < kTopoCap; ++b)
      if ((jk[s] >> b) & 1ull) r.ids.push_back((int64_t)(uint32_t)key[b]);
    std::sort(r.ids.begin(), r.ids.end());
    r.nt = cnt[s];
    for (int k = 0; k < 3; ++k) r.p[k] = (double)(long long)jp[3 * s + k] / kPosScale / cnt[s];
    recs.push_back(r);
  }
  std::sort(recs.begin(), recs.end(), [](const Rec& a, const Rec& b) { return a.ids < b.ids; });
  *n = (int32_t)recs.size();
  if (!nids) return GG_OK;
  if ((int32_t)recs.size() > cap || !ids8 || !ntri || !pos) return fail(c, GG_ERR_ARG, "destination too small");
  for (size_t r = 0; r < recs.size(); ++r) {
    nids[r] = (int32_t)recs[r].ids.size();
    for (int k = 0; k < 8; ++k) ids8[8 * r + k] = k < (int)recs[r].ids.size() ? (int32_t)recs[r].ids[k] : -1;
    ntri[r] = recs[r].nt;
    for (int k = 0; k < 3; ++k) pos[3 * r + k] = recs[r].p[k];
  }
  return GG_OK;
}

int gg_get_all_bounds(gg_ctx* c, int32_t which, int32_t* bounds) {
  if (!c || !bounds) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const DescDev& d = which ? dout(c) : din(c);
  HIPCHK(c, hipMemcpy(bounds, d.bounds, 24 * (size_t)(c->p.n_grains + 1), hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_download_all_lists(gg_ctx* c, int32_t which, int32_t* offsets, int32_t* ids, int64_t cap, int64_t* n) {
  if (!c || !offsets || !n) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  const int N = c->p.n_grains, M = c->p.max_list;
  std::vector<int32_t> sel(N + 1), len0(N + 1), len1(N + 1);
  HIPCHK(c, hipMemcpy(sel.data(), c->LD.sel, 4 * (size_t)(N + 1), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(len0.data(), c->LD.len[0], 4 * (size_t)(N + 1), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(len1.data(), c->LD.len[1], 4 * (size_t)(N + 1), hipMemcpyDeviceToHost));
  offsets[0] = 0;
  offsets[1] = 0;
  int64_t total = 0;
  for (int g = 1; g <= N; ++g) {
    const int b = which ? sel[g] ^ 1 : sel[g];
    total += b ? len1[g] : len0[g];
    offsets[g + 1] = (int32_t)total;
  }
  *n = total;
  if (!ids) return GG_OK;
  if (total > cap) return fail(c, GG_ERR_ARG, "destination too small");
  std::vector<int32_t> buf((size_t)(N + 1) * M);
  for (int b = 0; b < 2; ++b) {
    HIPCHK(c, hipMemcpy(buf.data(), c->LD.buf[b], 4 * buf.size(), hipMemcpyDeviceToHost));
    for (int g = 1; g <= N; ++g) {
      const int bg = which ? sel[g] ^ 1 : sel[g];
      if (bg != b) continue;
      std::copy(buf.begin() + (size_t)g * M, buf.begin() + (size_t)g * M + (offsets[g + 1] - offsets[g]),
                ids + offsets[g]);
    }
  }
  return GG_OK;
}

int gg_get_new_bounds(gg_ctx* c, int32_t* bounds) {
  if (!c || !bounds) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(bounds, c->d_newb, 24 * (c->p.n_grains + 1), hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_box_voxels(gg_ctx* c, int64_t* voxels) {
  if (!c || !voxels) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpy(voxels, c->d_voxels, 8, hipMemcpyDeviceToHost));
  return GG_OK;
}

int gg_enable_stage_timing(gg_ctx* c, int32_t enable) {
  if (!c) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  c->timing = enable != 0;
  return GG_OK;
}

int gg_get_stage_times(gg_ctx* c, double* ms, int64_t* launches, int32_t n, int32_t reset) {
  if (!c || !ms || n < 0) return GG_ERR_ARG;
  int rc = gg_synchronize(c);
  if (rc) return rc;
  for (int i = 0; i < n && i < kNumSeg; ++i) {
    ms[i] = c->seg_ms[i];
    if (launches) launches[i] = c->seg_n[i];
  }
  if (reset)
    for (int i = 0; i < kNumSeg; ++i) {
      c->seg_ms[i] = 0;
      c->seg_n[i] = 0;
    }
  return GG_OK;
}

}  // extern "C"
