// Internal declarations shared by the HIP translation units of libgg_levelset.
//
// Device data model (one context = every grain box of one simulation):
//  * Two distance-buffer roles, "in" and "out" (the reference's
//    m_inputDistance / m_outputDistance of every LSbox).  Each role is a
//    DescDev: one arena holding all grains' boxes back to back plus per-grain
//    bounds rows and arena offsets.  gg_switch_buffers swaps the roles on the
//    host (no data moves).
//  * Two IDLocal arenas with their own per-grain layout: the comparison reads
//    the previous step's IDLocal (for voxels no neighbour covers) and writes
//    the current one.
//  * Two comparison-list buffers; per grain a selector says which one holds
//    m_secondOrderNeighbours, the other receives m_comparisonList.
//  * A work plan: alive grains grouped by box-size class (side <= 32, 48, 64,
//    96) with per-class tile prefixes, so each stage runs one launch per class
//    with LDS sized for that class, and every launch reads its tile count from
//    device memory (no host synchronisation inside a timestep).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gg {

constexpr int kClasses = 4;
constexpr int kMaxN = 96;
constexpr int kColBlock = 64;      // columns per z-sweep / z-convolution tile
constexpr int kSpecBlock = 32;     // spectral columns per FP64 z-DFT tile

__host__ __device__ inline int class_max(int c) { return c == 0 ? 32 : c == 1 ? 48 : c == 2 ? 64 : 96; }
__host__ __device__ inline int size_class(int n) {
  return n <= 32 ? 0 : n <= 48 ? 1 : n <= 64 ? 2 : n <= 96 ? 3 : -1;
}

enum { kDeleted = 0, kAlive = 1, kVanished = 2, kDoomed = 3 };
// kDoomed lives only inside a level set: an alive grain without a voxel >= 0
// (it has no hull, so it vanishes), marked before the output plan so that the
// plan leaves it out; the level set's last part makes it vanished.

// device error bits
enum {
  kErrArena = 1,      // distance / IDLocal arena capacity exceeded
  kErrList = 2,       // comparison list capacity exceeded
  kErrBoxSize = 4,    // box side above kMaxN or box not cubic
  kErrSpec = 8,       // spectral scratch capacity exceeded
  kErrEmptyIO = 16,   // redistancing with an empty in/out intersection
  kErrHistory = 32,   // count history full
  kErrTopo = 64,      // a grain's hull touches more grains than its topology table holds
  kErrTri = 128,      // the correction's triangle pool is full
  kErrCmpWork = 256   // the comparison's spatial work list is full
};

// Hull topology (optional output of the level set): per grain, the grains its
// marching-cubes hull touches, as an open-addressed table of kTopoCap slots:
// key = grain id (-1 empty), ord = smallest first-appearance key of that id
// (cell index in the (z, y, x) loop * 8 + rank of the id among the cell's
// ascending ids).
// Junction lists: each cut cell's list of distinct ids, as a mask over the
// slots those ids hold in the grain's table; per grain an open-addressed set of
// kJuncCap masks (0 empty), counted by popcount on the host
// (GrainHull::computeGrainBoundaryElements' classes).
constexpr int kTopoCap = 64;
constexpr int kJuncCap = 512;  // ~160 lists at most per hull at 8000 grains
// With the convolution correction on, the hull also records every triangle's
// barycenter with the grain and the junction slot of its cell's list (an
// unsorted pool, appended to per tile), and each list's first appearance (the
// smallest cell index in the (z, y, x) loop whose list it is).
struct TriRec {
  double x, y, z;  // barycenter, grid coordinates
  int32_t g, slot;
};
constexpr double kPosScale = 4294967296.0;  // 2^32: fixed-point units of a junction position sum per grid unit
struct TopoDev {
  int32_t* key;                // [(N+1) * kTopoCap]
  int32_t* ord;                // [(N+1) * kTopoCap]
  unsigned long long* jkey;    // [(N+1) * kJuncCap]
  int on;                      // 0: not computed
  unsigned long long* jpos;    // [(N+1) * kJuncCap * 3] barycenter sums of each list's triangles, fixed point
                               // (kPosScale units per grid unit), so the sums do not depend on the order of the adds
  int32_t* jcnt;               // [(N+1) * kJuncCap] triangles of each list
  double* jarea;               // [(N+1) * kJuncCap] summed triangle areas of each list (grid units)
  int corr;                    // also the triangle records and first appearances below
  int32_t* jord;               // [(N+1) * kJuncCap] first appearance of each list
  TriRec* raw;                 // [raw_cap] triangle records of the last hull
  unsigned long long* raw_n;   // records written (may exceed raw_cap: then the topology error bit is set)
  int64_t raw_cap;
};

// Interfacial elements of every grain's last hull and the convolution
// correction's inputs (src/box.cpp:428-536, src/grainHull.cpp:187-274,
// 483-531, 570-660).  Elements are kept per grain in the order
// projectPointToGrainBoundary visits them: high-order junctions, quadruple
// junctions, triple lines, boundaries.  An element's neighbour set is a mask
// over the grain's hull-id table slots (TopoDev::key).
constexpr int kElemCap = kJuncCap;
constexpr int kChunk = 16;         // barycenters per bounding box of the nearest-element search
constexpr double kFarPoint = 1e30;  // coordinate of a chunk's padding points
struct CorrDev {
  int32_t* ready;              // set by the element build: a hull with elements exists (run_sim's loop != 0)
  int see_on;                  // Settings::UseStoredElasticEnergy
  int aniso;                   // energies / mobilities from the pair tables (else 1)
  double tl_inv;               // 1 / (3 sqrt(3 ds^2) * TripleLineDrag), TripleLine::computeMobility
  double dt;
  double tsq;                  // TimeSlope^2
  const double* gam;           // [(N+1)^2] Read-Shockley energy of the ordered pair (a, b)
  const double* mob;           // [(N+1)^2] mobility of the ordered pair
  const double* see;           // [N+1] stored elastic energy (0: the wall, pinned 0)
  int32_t* ecount;             // [N+1] elements of the grain
  int32_t* eclass;             // [(N+1) * 4] high-order junctions, quadruple junctions, triple lines (rest: boundaries)
  double* epos;                // [(N+1) * kElemCap * 3] junction positions (0 for lines and boundaries)
  unsigned long long* emask;   // [(N+1) * kElemCap] neighbour set
  double* energy;              // [(N+1) * kElemCap] m_energy
  double* mobility;            // [(N+1) * kElemCap] m_mobility
  int32_t* efirst;             // [(N+1) * (kElemCap + 1)] first barycenter of each element, relative to tbase
  int32_t* ecur;               // [(N+1) * kElemCap] fill cursors
  int32_t* slot_elem;          // [(N+1) * kJuncCap] element of each junction slot
  int64_t* tbase;              // [N+1] the grain's first barycenter in tri
  double* tri;                 // [3 * tri_cap] barycenters grouped by element, each element padded
                               // to whole chunks of kChunk with far-away points
  float* cbox;                 // [6 * (tri_cap / kChunk)] bounding box of each chunk (lo xyz, hi xyz)
  unsigned long long* tri_n;   // barycenters placed
  int64_t tri_cap;
};

struct DescDev {
  void* arena;       // double* (FP64) or float* (FP32)
  int32_t* bounds;   // [N+1][6] {x0,y0,z0,x1,y1,z1}
  int64_t* off;      // [N+1] offset of the box in the arena (elements)
};

struct IdlDev {
  uint32_t* arena;
  int32_t* bounds;
  int64_t* off;
};

struct ListDev {
  int32_t* buf[2];   // [(N+1)*max_list]
  int32_t* len[2];   // [N+1]
  int32_t* sel;      // [N+1] buffer index holding m_secondOrderNeighbours
  int32_t max_list;
  // Neighbour table of the comparison, built at the start of the comparison
  // stage from m_comparisonList: per grain, in list order, the entries whose
  // input box overlaps the grain's box, each as
  //   x, y: B (48 bits: arena index of the shared rectangle's first voxel in the
  //         neighbour's layer q2), y >> 16: the neighbour's box side nn
  //   z: the shared x/y rectangle, local to the grain's box: x0 | x1<<8 | y0<<16 | y1<<24
  //   w: the neighbour's z range q2 | q5 << 16
  // so the rectangle's first voxel in layer k is B + (k - q2) * nn^2.
  int4* tab;         // [(N+1)*max_list]
  int32_t* tab_id;   // [(N+1)*max_list] neighbour id of each entry
  int32_t* tab_len;  // [N+1]
};

// Tile quantities carried per class by the plan.
enum { kQLayers = 0, kQCols = 1, kQSpecCols = 2, kNumQ = 3 };

struct PlanDev {
  int32_t* cls_count;   // [kClasses]
  int32_t* cls_start;   // [kClasses] first index into grains[]
  int32_t* grains;      // [N] alive grain ids grouped by class, in `order` inside a class
  int32_t* grain_pos;   // [N+1] index of a grain in grains[] (-1 if not alive)
  int32_t* order;       // [N] permutation of 1..N: the spatial (Morton) order of the initial box centres
  int32_t* vorder;      // [N] work order of the step when it is volume-sorted (nullptr: the spatial order)
  int32_t* ncls;        // [N] per position in `order`: box side << 8 | size class, or -1 (not planned)
  int32_t* pre[kNumQ];  // [N + kClasses] per-class exclusive prefixes (segment c at cls_start[c]+c)
  int32_t* totals;      // [kNumQ * kClasses]
  int64_t* layer_base;  // [kClasses] first hull-partial slot of each class
  int32_t* tile_base;   // [kNumQ * kClasses] first entry of each class in tmap[q]
  uint32_t* tmap[kNumQ];  // [kMaxTilesPerGrain * N] tile -> grain << 8 | tile index inside the grain
};
// Spatial work list of the comparison: every wavefront tile of every planned
// grain (column block, run of layers), bucketed by the cell of its centre.
// Cells are 16 x 8 x 8 voxels (x, y, z); a key is the raster index of the
// cell's brick of 8 x 8 x 8 cells times 512 plus the cell's Morton index
// inside the brick, so consecutive keys are neighbouring cells.
struct CmpSpaceDev {
  int2* wl;      // [cap] (grain, column block | run << 16), in key order
  int32_t* cnt;   // [nkeys + 1] counts, then bucket starts (scan), then ends (fill); [nkeys] = tiles
  int32_t* bsum;  // [nkeys / 4096 + 1] block sums of the scan
  int nkeys, cap, bx, by;  // bricks per axis: bx (x), by (y and z)
};
constexpr int kCmpSpaceShapes = 4;  // (layers per run, gathers in flight): (8,8) (16,8) (16,16) (32,8)
constexpr int kCmpVariants = 8;  // comparison tile shapes the autotuner chooses from
constexpr int kCmpDefault = 1;   // 16 layers, 8 gathers in flight
constexpr int kMaxGrid = 32767;  // ngridpoints limit

// Tiles of one grain for any quantity: n layers <= 96, ceil(n*n/64) <= 144
// column blocks, ceil(n*(n/2+1)/32) <= 147 spectral blocks; tile indices fit 8
// bits and grain ids 24 bits of a tmap entry, and tile counts stay below 2^31.
constexpr int kMaxTilesPerGrain = 160;
constexpr int kMaxGrains = 8000000;

struct SimParams {
  int32_t N;
  int32_t ngp;
  int32_t gb;
  int32_t tube;
  double h;
  double delta;
};

// Per-n tables (n = 1..kMaxN), host-computed in FP64.
struct TablesDev {
  const int32_t* tw_off;  // [kMaxN+1] offset into twc/tws and circ
  const double* twc;      // cos(2 pi t / n)
  const double* tws;      // sin(2 pi t / n)
  const int32_t* g_off;   // [kMaxN+1] offset into gtab
  const double* gtab;     // spectral factor G(s), s = kx2^2+ky2^2+kz2^2
  const float* circ;      // FP32 circulant heat-kernel row (one axis) per n
  const int32_t* fplan;   // [(kMaxN+1) * 8] FFT radices per n in pass order, 0-terminated
};

// ---- helpers --------------------------------------------------------------
__device__ inline int box_n(const int32_t* b, int g) { return b[6 * g + 3] - b[6 * g]; }

__device__ inline void cube_bounds(int* b, int maximumLength) {  // resizeToCube, src/dimensionalBuffer.h:139-219
  const int grid_size = maximumLength - 1;
  const int w = b[3] - b[0], hgt = b[4] - b[1], d = b[5] - b[2];
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

// Locate tile t of class c for quantity q: grain id and local tile index.
__device__ inline void plan_tile(const PlanDev& P, int c, int q, int t, int& g, int& local) {
  const uint32_t v = P.tmap[q][P.tile_base[q * kClasses + c] + t];
  g = (int)(v >> 8);
  local = (int)(v & 0xFFu);
}

// Tiles of a grid-stride loop, XCD-aware.  The GPU hands workgroups to its 8
// XCDs round-robin (workgroup b runs on XCD b % 8) and every XCD has its own
// L2, so XCD x takes the contiguous tile range [x*T/8, (x+1)*T/8): the tiles of
// neighbouring grains (the plan orders grains spatially) then share one L2.
struct TileRange {
  int first, end, step;
};
constexpr int kXcds = 8;
__device__ inline TileRange xcd_tiles(int total) {
  const int G = gridDim.x, b = blockIdx.x;
  if (G % kXcds) return TileRange{b, total, G};
  const int x = b % kXcds, r = b / kXcds;
  const int lo = (int)((int64_t)total * x / kXcds), hi = (int)((int64_t)total * (x + 1) / kXcds);
  return TileRange{lo + r, hi, G / kXcds};
}

// Flattened tile lookup over all classes for one quantity.
__device__ inline bool plan_tile_any(const PlanDev& P, int q, int t, int& c, int& g, int& local) {
  for (c = 0; c < kClasses; ++c) {
    int tot = P.totals[q * kClasses + c];
    if (t < tot) {
      plan_tile(P, c, q, t, g, local);
      return true;
    }
    t -= tot;
  }
  return false;
}

__device__ inline int plan_total(const PlanDev& P, int q) {
  int s = 0;
  for (int c = 0; c < kClasses; ++c) s += P.totals[q * kClasses + c];
  return s;
}

template <class T>
__device__ inline int sgnv(T v) { return (T(0) < v) - (v < T(0)); }

// Copy count elements from global memory to LDS with kStageBatch loads in
// flight per lane before their LDS writes (a plain element loop waits for each
// load before its write).
constexpr int kStageBatch = 8;
template <class TD, class TS>
__device__ inline void stage_to_lds(TD* __restrict__ lds, const TS* __restrict__ src, int count, int tid, int nt) {
  for (int base = 0; base < count; base += nt * kStageBatch) {
    TS v[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int e = base + u * nt + tid;
      if (e < count) v[u] = src[e];
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int e = base + u * nt + tid;
      if (e < count) lds[e] = (TD)v[u];
    }
  }
}

// Same for a rows x width block whose rows are `stride` elements apart in
// global memory; element (r, c) goes to lds[r * lds_stride + c].
template <class TD, class TS>
__device__ inline void stage_rows_to_lds(TD* __restrict__ lds, int lds_stride, const TS* __restrict__ src,
                                         int64_t stride, int rows, int width, int tid, int nt) {
  const int count = rows * width;
  for (int base = 0; base < count; base += nt * kStageBatch) {
    TS v[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int e = base + u * nt + tid;
      if (e < count) {
        const int r = e / width, c = e - r * width;
        v[u] = src[(int64_t)r * stride + c];
      }
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int e = base + u * nt + tid;
      if (e < count) {
        const int r = e / width, c = e - r * width;
        lds[r * lds_stride + c] = (TD)v[u];
      }
    }
  }
}

__device__ inline int64_t lin3(int n, int z, int y, int x) { return ((int64_t)z * n + y) * n + x; }

}  // namespace gg

// ---- host-side launch wrappers (defined in the .hip files) -----------------
struct gg_ctx;
namespace gg {
struct Launch {
  hipStream_t stream;
  int grid;   // workgroups per grid-stride launch (8 per compute unit)
  int cus;    // compute units of the device
};
// Workgroups of a tile launch: per_cu per compute unit.  Grids of about one
// workgroup per tile let the hardware dispatcher balance tiles of uneven cost
// (a new workgroup starts wherever one finishes), which a persistent grid with
// a fixed share of tiles per workgroup does not; the grid-stride loop of each
// kernel covers tiles beyond the grid.  The values at the launch sites are the
// fastest of a sweep on the 8000-grain workload.
// Classes 2 and 3 (boxes of side 49..96) hold few grains in the workloads
// measured here; their launches get 1/16 of that (at least 8 per CU), so a
// nearly empty class does not pay for dispatching a full-size grid.
inline int tile_grid(const Launch& L, int per_cu, int cls) {
  return L.cus * (cls >= 2 ? (per_cu / 16 > 8 ? per_cu / 16 : 8) : per_cu);
}

void launch_plan(const Launch& L, const SimParams& sp, const int32_t* state, const int32_t* bounds,
                 int64_t* off_out, int64_t* spec_off, const PlanDev& P, int64_t arena_cap,
                 int64_t spec_cap, int32_t* err, int64_t* voxels);
// Grid layers [k_lo, k_hi) of the initial field; idfield holds those layers.
void launch_init_phi(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din,
                     const int32_t* idfield, bool fp32, int k_lo, int k_hi);
// part 0: (x,y) forward DFT | x,y circulant; part 1: z; part 2: (y,x) inverse DFT (FP64 only)
// convolution forms: the FP64 FFT, the FP32 circulant, the split-BF16 circulant
enum { kConvFFT64 = 0, kConvCirculantF32 = 1, kConvSplitBF16 = 2 };
void launch_conv(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                 const DescDev& dout, const TablesDev& tb, const int64_t* spec_off, void* spec, int mode);
void launch_idl_resize(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& dout,
                       const IdlDev& il);
void launch_init_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& d,
                       const ListDev& ld, int32_t* err);
void launch_lists(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                  const DescDev& dout, const ListDev& ld, int32_t* newb, int32_t* err);
// coarsen: the table is built from m_secondOrderNeighbours (the lists of the
// last comparison, as recalculateIDLocal's comparison reads them) and passes
// over neighbours that are not alive.
void launch_comparison_table(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                             const ListDev& ld, bool coarsen = false);
// variant: tile shape (0..kCmpVariants-1), chosen per size class by the autotuner
void launch_comparison(const Launch& L, int cls, int variant, const SimParams& sp, const PlanDev& P,
                       const DescDev& din, const DescDev& dout, const ListDev& ld, const IdlDev& il_old,
                       const IdlDev& il_new, int32_t* newb, bool fp32);
// All classes in one launch, wavefront tiles in spatial order (the work list W
// is rebuilt from the plan first); shape 0..kCmpSpaceShapes-1.  part: both
// steps, only the work-list build (it reads the plan and din's bounds), or
// only the comparison over a work list built for the same shape.
enum { kSpaceAll = 0, kSpaceBuild = 1, kSpaceRun = 2 };
void launch_comparison_space(const Launch& L, int shape, const SimParams& sp, const PlanDev& P, const DescDev& din,
                             const DescDev& dout, const ListDev& ld, const IdlDev& il_old, const IdlDev& il_new,
                             int32_t* newb, const CmpSpaceDev& W, int32_t* err, bool fp32, int part = kSpaceAll);
// Clears the grains' topology tables; before every hull with topology on.
void launch_topo_clear(const Launch& L, const SimParams& sp, const TopoDev& topo);
void launch_hull(const Launch& L, const SimParams& sp, const PlanDev& P, const DescDev& din, int32_t* exflag,
                 double* partials, const IdlDev& il, const TopoDev& topo, int32_t* err, bool fp32, int c0 = 0,
                 int c1 = kClasses);
// Parts of the level-set finalize (k_finalize).
constexpr int kFinBounds = 1, kFinState = 2, kFinCoarsen = 4;
void launch_finalize(const Launch& L, const SimParams& sp, const PlanDev& P, int32_t* state,
                     const ListDev& ld, int32_t* exflag, const double* partials, const int32_t* newb,
                     const DescDev& din, const DescDev& dout, double* volume, double* area,
                     int mode = kFinBounds | kFinState);
// Grid coarsening (LSbox::resizeGrid): up = h_old/h_new, down = h_new/h_old,
// delta = the new delta; sp holds the new grid.
struct CoarsenArgs {
  int32_t gb;
  double up, down, delta;
};
// New bounds of every alive grain: src bounds scaled (-> nb), cubed (-> dst).
void launch_coarsen_bounds(const Launch& L, const SimParams& sp, const CoarsenArgs& a, const int32_t* state,
                           const int32_t* src_bounds, int32_t* dst_bounds, int32_t* nb);
// Trilinear resampling of src (old grid) into dst (new bounds and layout).
void launch_coarsen_resample(const Launch& L, const SimParams& sp, const CoarsenArgs& a, const int32_t* state,
                             const DescDev& src, const DescDev& dst, bool fp32);
// The size^3 voxel container (grain id per voxel, 0 for none) of the buffers d.
void launch_container(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& d, uint32_t* cont,
                      int size, bool fp32);
// part 0: x/y sweeps; part 1: z sweeps and clamp
void launch_redist(const Launch& L, int part, const SimParams& sp, const PlanDev& P, const DescDev& din,
                   const DescDev& dout, int32_t* err, bool fp32, int c0 = 0,
                   int c1 = kClasses);
void launch_post_redist(const Launch& L, const SimParams& sp, const int32_t* state, const DescDev& din,
                        const DescDev& dout);
void launch_count(const Launch& L, const SimParams& sp, const int32_t* state, int32_t* step,
                  int32_t* history, int32_t history_cap, int32_t* err);
void upload_mc_tables();
// Convolution correction (gg_correct.hip): the elements of every alive grain
// from its last hull (after launch_hull with topo.corr), their barycenters
// grouped by element, and the correction of the convolved output buffer.
void launch_elements(const Launch& L, const SimParams& sp, const int32_t* state, const TopoDev& topo,
                     const CorrDev& cd, int32_t* err);
void launch_correction(const Launch& L, const SimParams& sp, const PlanDev& P, const int32_t* state,
                       const DescDev& din, const DescDev& dout, const IdlDev& il_stale, const TopoDev& topo,
                       const CorrDev& cd, bool fp32);
}  // namespace gg
