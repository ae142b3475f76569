/*
 * gg_levelset.h -- C ABI of the MI355X per-grain level-set timestep library.
 *
 * One context owns every grain box of one simulation on one GPU: the two
 * distance-buffer arenas (the reference's m_inputDistance / m_outputDistance
 * of every LSbox), the IDLocal arenas, the comparison lists, the spectral
 * scratch and the per-size-class work plans.  Each stage entry point runs all
 * grains in one batched launch per size class on the context's HIP stream.
 *
 * The entry points replace the LSbox operator surface that grainhdl's OpenMP
 * loops call (paths relative to the GraGLeS_3D source tree):
 *
 *   gg_convolution             LSbox::executeConvolution        src/box.h:172, src/box.cpp:414-539
 *   gg_switch_buffers          LSbox::switchInNOut              src/box.h:190, src/box.cpp:803-808
 *                              (grainhdl::switchDistancebuffer  src/grainhdl.cpp:1166-1178)
 *   gg_update_neighbour_lists  LSbox::computeSecondOrderNeighbours src/box.h:129, src/box.cpp:1090-1134
 *   gg_comparison              LSbox::executeComparison + boundaryCondition + executeSetComparison
 *                              src/box.h:123,192,128; src/box.cpp:891-966, 997-1088, 814-878
 *                              (grainhdl::comparison_box        src/grainhdl.cpp:654-667)
 *   gg_level_set               grainhdl::level_set -> LSbox::extractContour
 *                              src/grainhdl.cpp:669-684, src/box.cpp:1144-1161 (exists, box
 *                              resize + resizeToCube, grain volume and surface area)
 *   gg_redistancing            LSbox::executeRedistancing       src/box.h:117, src/box.cpp:2322-2586
 *   gg_count_grains            grainhdl::countGrains            src/grainhdl.cpp:936-943
 *   gg_timestep                one run_sim iteration            src/grainhdl.cpp:836-919
 *
 * Conventions: grain ids are 1..n_grains, id 0 is the domain wall.  A bounds
 * row is {xmin, ymin, zmin, xmax, ymax, zmax}, half-open, in global grid
 * points; box storage is x fastest, then y, then z (src/dimensionalBuffer.h:75).
 * Every call returns GG_OK (0) or an error code; gg_last_error() describes the
 * last failure.  No call throws.  A context is used by one host thread.
 */
#ifndef GG_LEVELSET_H
#define GG_LEVELSET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_ctx gg_ctx;

enum gg_status {
  GG_OK = 0,
  GG_ERR_ARG = 1,      /* invalid argument */
  GG_ERR_HIP = 2,      /* a HIP runtime call failed */
  GG_ERR_CAPACITY = 3, /* arena, list or box-size capacity exceeded on the device */
  GG_ERR_STATE = 4,    /* call out of order (e.g. no boxes set) */
  GG_ERR_DEVICE = 5    /* no usable GPU */
};

enum gg_precision {
  GG_FP64_PLAIN = 0,  /* all FP64; bit-identical to the FP64 CPU restatement */
  GG_FP32_BATCHED = 1 /* FP32 distance fields, FP64 hull reductions */
};

enum gg_grain_state { GG_DELETED = 0, GG_ALIVE = 1, GG_VANISHED = 2 };

/* Grid and time variables of grainhdl (src/grainhdl.cpp:1419-1444, 237-239). */
typedef struct gg_params {
  int32_t ngridpoints;         /* grainhdl::ngridpoints = realDomainSize + 2*grid_blowup */
  int32_t grid_blowup;         /* grainhdl::grid_blowup = DomainBorderSize */
  int32_t conv_domain_size;    /* realDomainSize as read by the convolution (rds-1 in Voronoi mode) */
  int32_t boundary_grain_tube; /* grainhdl::BoundaryGrainTube; never initialised by the reference, pass 0 */
  double h;                    /* grainhdl::h = 1/realDomainSize */
  double delta;                /* grainhdl::delta = DomainBorderSize/realDomainSize */
  double dt;                   /* grainhdl::dt = TimeStepSize/realDomainSize^2 */
  int32_t precision;           /* enum gg_precision */
  int32_t n_grains;            /* Settings::NumberOfParticles (at most 8000000) */
  int32_t max_list;            /* capacity of one comparison list (0 = 512, at most 1024) */
  int32_t device;              /* HIP device ordinal */
  int64_t arena_voxels;        /* voxels per distance arena (0 = 2x the initial sum of box volumes) */
} gg_params;

/* Create / destroy a context.  Replaces the per-grain buffer allocation of the
 * LSbox constructors (src/box.cpp:111-119, 213-221) and grainhdl's per-thread
 * FFT memory pools (src/grainhdl.cpp:598-609). */
int gg_create(const gg_params* params, gg_ctx** out);
int gg_destroy(gg_ctx* ctx);
const char* gg_last_error(const gg_ctx* ctx);

/* Box bounds of every grain as the LSbox constructor computes them before
 * resizeToCube (src/box.cpp:185-221): rows 0..n_grains of 6 int32 (row 0
 * ignored), exists[id] != 0 for grains that have a box.  The library applies
 * resizeToCube(ngridpoints) to both buffers and sizes IDLocal to them.  The
 * first call sizes the device arenas (params.arena_voxels, or twice the boxes'
 * voxels plus 65536); a later call whose boxes need more voxels than that
 * returns GG_ERR_CAPACITY. */
int gg_set_boxes(gg_ctx* ctx, const int32_t* bounds, const uint8_t* exists);

/* calculateDistanceFunction + executeRedistancing of every grain from the
 * voxelised ID field (ngridpoints^3 int32, x fastest), src/box.cpp:374-397,
 * driven by grainhdl::distanceInitialisation (src/grainhdl.cpp:580-597). */
int gg_init_distance(gg_ctx* ctx, const int32_t* idfield);

/* gg_init_distance without any device allocation (it has no counterpart in the
 * reference): for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The ID
 * field is staged from idfield in runs of grid layers through a 16 MiB device
 * array per GPU that the library holds until the process ends; the result is
 * gg_init_distance's.  Needs ngridpoints^2 <= 4194304 (ngridpoints <= 2048),
 * else GG_ERR_CAPACITY.  Every context on one GPU shares that GPU's array,
 * which keeps the last call's contents: calls on one GPU are serialised by a
 * host lock (each returns synchronised), so they never overlap. */
int gg_init_distance_noalloc(gg_ctx* ctx, const int32_t* idfield);

/* Initial m_secondOrderNeighbours of every grain (grainhdl::find_neighbors,
 * src/grainhdl.cpp:1029-1054) as CSR: offsets[0..n_grains+1], ids. */
int gg_set_neighbour_lists(gg_ctx* ctx, const int32_t* offsets, const int32_t* ids);

/* The initial m_secondOrderNeighbours built on the device from the boxes given
 * to gg_set_boxes (grainhdl::find_neighbors, src/grainhdl.cpp:1029-1054, and
 * LSbox::computeDirectNeighbours, src/box.cpp:2871-2887): for every alive grain
 * the other alive grains whose cubed boxes overlap it with inclusive bounds,
 * in ascending id (the R-tree's traversal order is not reproduced) -- the same
 * lists gg_initial_neighbour_lists gives on the host.  Replaces
 * gg_set_neighbour_lists at set-up.  Tests all pairs (a one-off; about 10 ms at
 * 50 000 grains). */
int gg_build_neighbour_lists(gg_ctx* ctx);

/* Stage entry points (see the table above).  Order: gg_redistancing writes
 * the output boxes that the preceding gg_level_set laid out, so it follows a
 * gg_level_set (or gg_init_distance's own); called again before the next
 * gg_level_set, or after gg_coarsen_grid, it returns GG_ERR_STATE. */
int gg_switch_buffers(gg_ctx* ctx);
int gg_convolution(gg_ctx* ctx);
int gg_update_neighbour_lists(gg_ctx* ctx);
int gg_comparison(gg_ctx* ctx);
int gg_level_set(gg_ctx* ctx);
int gg_redistancing(gg_ctx* ctx);

/* The coarsening branch of grainhdl::gridCoarsement (src/grainhdl.cpp:1446-1492),
 * which a coarsening step runs in place of the opening switch: the grid and
 * time variables of the new size (updateGridAndTimeVariables,
 * src/grainhdl.cpp:1416-1444: ngridpoints = new_size + 2*grid_blowup,
 * h = 1/new_size, delta = grid_blowup/new_size, the convolution's domain size
 * new_size; dt = new_dt, which the caller computes for its convolution mode),
 * then per grain LSbox::resizeGrid (src/box.cpp:2595-2697, trilinear
 * resampling of the last redistancing's result), recalculateIDLocal
 * (src/box.cpp:2699-2702) and extractContour (src/box.cpp:1144-1161).  A
 * coarsening step is gg_coarsen_grid followed by gg_timestep: the call ends
 * with the switch that gg_timestep's opening switch undoes.  Grains that
 * vanished at the level set before it stay vanished and the comparison passes
 * over them (the reference resamples their pre-comparison field and runs their
 * hull again).  1 <= new_size <= ngridpoints - 2*grid_blowup. */
int gg_coarsen_grid(gg_ctx* ctx, int32_t new_size, double new_dt);

/* The voxel container of grainhdl::saveNetworkAsVoxelContainer
 * (src/grainhdl.cpp:1074-1154, the Container_size_<size>_t_<loop>.raw data)
 * with LSbox::copyDataToContainer (src/box.cpp:3008-3032): size^3 uint32
 * (x fastest), size = realDomainSize as the run holds it (the convolution's
 * domain size); a voxel holds the id of a grain whose output buffer is >= 0
 * there, shifted by -grid_blowup, else 0.  Call it after a whole timestep.
 * Where several grains qualify the largest id stays (the reference's write
 * order with one thread); only alive grains are written (the reference also
 * writes a grain that vanished in that step, from its pre-comparison field,
 * which the device does not keep).  *size is set first; dst may be NULL to
 * query it; cap is in elements. */
int gg_voxel_container(gg_ctx* ctx, uint32_t* dst, int64_t cap, int32_t* size);

/* The interfacial elements of grain id's hull as of the last level set with
 * topology on (GrainHull::computeGrainBoundaryElements and
 * subDivideTrianglesToInterfacialElements, src/grainHull.cpp:187-270; the
 * position of QuadrupleJunction / HighOrderJunction::computePosition,
 * src/InterfacialElement.cpp:406-469): one record per distinct cell id list,
 * sorted by its ids: nids[r] ids ascending in ids8[8r..8r+7] (-1 padded, the
 * first 8 of longer lists), ntri[r] triangles, pos[3r..3r+2] the mean of their
 * barycenters in grid coordinates (each cell's barycenter sum rounded to 2^-32
 * grid units and summed as integers, so the result does not depend on the
 * order of the sums and is within 2^-33 of the serial sum).  Junctions are not
 * merged (mergeJunction).  *n is set first; nids may be NULL to query it. */
int gg_download_hull_elements(gg_ctx* ctx, int32_t id, int32_t* nids, int32_t* ids8, int32_t* ntri, double* pos,
                              int32_t cap, int32_t* n);

/* ---- convolution correction (configurations with
 * Settings::DisableConvolutionCorrection = 0) ---- */

/* Settings of the correction branch of LSbox::executeConvolution
 * (src/box.cpp:428-536; the magnetic term is not part of this path). */
typedef struct gg_physics {
  int32_t correction;       /* !Settings::DisableConvolutionCorrection */
  int32_t mobility_model;   /* Settings::UseMobilityModel: sigmoidal GB mobility in the misorientation */
  int32_t unique_energies;  /* Settings::UniqueGBEnergies: 1 = unit GB energy, 0 = Read-Shockley */
  int32_t stored_energy;    /* Settings::UseStoredElasticEnergy */
  double triple_line_drag;  /* Settings::TripleLineDrag (0 gives triple lines zero mobility) */
} gg_physics;

/* Grain counts up to which misorientation-dependent energies or mobilities are
 * accepted: the ordered-pair tables take 16 (n_grains+1)^2 bytes of device
 * memory. */
#define GG_MAX_ANISO_GRAINS 16384

/* Turns the convolution correction on (or off) for the following steps.
 * After every hull (gg_level_set, gg_timestep, gg_coarsen_grid) the grain's
 * interfacial elements are built as GrainHull::computeGrainBoundaryElements,
 * subDivideTrianglesToInterfacialElements and computeJunctionPosition with
 * mergeJunction build them (src/grainHull.cpp:187-274, 483-531), with the
 * energies and mobilities of src/InterfacialElement.cpp:52-404 (grain
 * boundaries: Read-Shockley energy and misorientation mobility of the pair;
 * triple lines: sigma = g(me,a) - g(a,b) + g(me,b) clamped to [0.01, 1],
 * mobility 1/(1/(3 sqrt(3 ds^2) TripleLineDrag) + mean pair mobility), ds =
 * 1/conv_domain_size; quadruple junctions: means over their three triple
 * lines; high-order junctions: 1, 1); after every later convolution each
 * voxel with |phi_in| < delta is corrected with the element nearest to it
 * among those bordering its IDLocal grain (projectPointToGrainBoundary,
 * src/grainHull.cpp:570-660): phi_out = phi_in + (phi_out - phi_in) * m * g,
 * then, with stored energies, minus (see[self] - see[id]) * g * dt /
 * TimeSlope^2 (TimeSlope = 1/1.21) for an id that is the wall or a grain not
 * deleted -- where (m, g) are the element's mobility and energy (GBInfo's
 * swapped fields, as get_GBInfo passes them).  The first convolution after the
 * call is not corrected (run_sim's loop 0, before any hull).
 * quat: 4 doubles per id 0..n_grains (id 0: the domain wall's pseudo grain),
 * needed when energies or mobilities depend on the misorientation; see:
 * n_grains+1 doubles (see[0] for the wall), needed with stored_energy.  The
 * pair tables are computed on the host (gg_gb_pair_tables).  Drops a
 * captured step graph; enables the hull topology. */
int gg_set_physics(gg_ctx* ctx, const gg_physics* phys, const double* quat, const double* see);

/* The interfacial elements of grain id's last hull, in the order the
 * projection visits them (high-order junctions, quadruple junctions, triple
 * lines, grain boundaries): per element its class (3, 2, 1, 0), the grains it
 * borders besides the grain itself (nnb[e], ascending in nb8[8e..8e+7], -1
 * padded), its triangles (nbary), position (junctions: the mean barycenter,
 * summed in no fixed order; else 0), energy and mobility.  *n is set first;
 * cls may be NULL to query it. */
int gg_download_elements(gg_ctx* ctx, int32_t id, int32_t* cls, int32_t* nnb, int32_t* nb8, int32_t* nbary,
                         double* pos, double* energy, double* mobility, int32_t cap, int32_t* n);

/* countGrains: synchronises and returns the number of grains not deleted. */
int gg_count_grains(gg_ctx* ctx, int32_t* count);

/* nsteps iterations of the run_sim body (switch, convolution, switch, lists,
 * comparison, switch, level set, redistancing, count).  use_graph != 0 replays
 * a captured HIP graph of two timesteps (captured on first use). */
int gg_timestep(gg_ctx* ctx, int32_t nsteps, int32_t use_graph);

/* Autotuner of the comparison kernel (no counterpart in the reference, whose
 * only tuning is its static grain schedulers, src/IterativeGrainScheduler.cpp,
 * src/SquaresGrainScheduler.cpp): runs one timestep, and inside it times each
 * of the comparison's tile shapes per box-size class (n <= 32, 48, 64, 96) on
 * the step's own inputs (reps launches each), keeps the fastest per class and
 * drops a captured step graph.  The shapes (z-layers per wavefront, gathers in
 * flight per neighbour) are 0: (8, 8), 1: (16, 8), 2: (16, 16), 3: (32, 8)
 * with a grain's tiles column-block-major, and 4..7 the same four run-major
 * (consecutive wavefronts take consecutive column blocks of one run of
 * layers); every shape gives the same result.  choice: 4 entries (may be
 * NULL); ms: 4 * GG_CMP_VARIANTS mean launch times (may be NULL).  With the
 * comparison in spatial order (gg_set_comparison_order) it times the four
 * shapes 0..3 of that order's single launch instead (work-list build
 * included) and keeps the fastest: every choice entry receives it, ms[0..3]
 * the four times. */
#define GG_CMP_VARIANTS 8
int gg_autotune(gg_ctx* ctx, int32_t reps, int32_t* choice, float* ms);
/* Order of the comparison's wavefront tiles (no counterpart in the reference,
 * whose comparison_box runs grain by grain, src/grainhdl.cpp:654-667).
 * GG_CMP_ORDER_GRAIN (default): one launch per size class, a grain's tiles
 * together (grains in the work order).  GG_CMP_ORDER_SPACE: one launch for
 * all classes whose tiles are sorted by the cell of their centre (16 x 8 x 8
 * voxels, cells in brick-raster then Morton order; the list is rebuilt on the
 * device at every comparison), so the tiles of the grains whose boxes cover a
 * voxel run together and read its neighbour values from one L2; tile shape
 * 0..3 as for gg_autotune (kept, but unused, in grain order).  Results do
 * not depend on the order or the shape.  gg_get_comparison_order also
 * returns the spatial order's shape (may be NULL). */
#define GG_CMP_ORDER_GRAIN 0
#define GG_CMP_ORDER_SPACE 1
int gg_set_comparison_order(gg_ctx* ctx, int32_t order, int32_t shape);
int gg_get_comparison_order(gg_ctx* ctx, int32_t* order, int32_t* shape);
/* The shape per size class in use (4 entries), and setting it (for instance a
 * choice saved from an earlier gg_autotune); default 1 for every class. */
int gg_get_tuning(gg_ctx* ctx, int32_t* variants);
/* Order in which every stage's launches take the grains (no counterpart in
 * the reference, whose grainhdl hands grains to OpenMP threads by static
 * schedulers, src/IterativeGrainScheduler.cpp:13-26,
 * src/SquaresGrainScheduler.cpp:26-105).  GG_ORDER_SPATIAL (default): the
 * Morton order of the initial box centres, so tiles that run together read
 * neighbouring boxes; GG_ORDER_VOLUME: a device-side queue rebuilt at every
 * plan, grains sorted by box side, largest first (equal sides in the spatial
 * order), so the largest boxes start first and the tail of a launch is made
 * of small tiles.  Results do not depend on the order. */
/* Form of the FP32 batched path's heat-kernel convolution (executeConvolution,
 * src/box.cpp:414-539, the MKL/FFTW transforms of :685-795).
 * GG_CONV_DEFAULT: the separable circulant in FP32 (x, y on the vector units,
 * z on the matrix cores; the FP64 path always uses its FFT).
 * GG_CONV_SPLIT_BF16: the three circulant passes as dense matrix products on
 * the matrix cores with every FP32 operand split into two BF16 parts and three
 * of the four partial products (all but lo*lo) accumulated in FP32: a relative
 * error of about 2^-16 per product.  GG_ERR_ARG on the FP64 path. */
#define GG_CONV_DEFAULT 0
#define GG_CONV_SPLIT_BF16 1
int gg_set_convolution_form(gg_ctx* ctx, int32_t form);
#define GG_ORDER_SPATIAL 0
#define GG_ORDER_VOLUME 1
int gg_set_work_order(gg_ctx* ctx, int32_t order);
int gg_get_work_order(gg_ctx* ctx, int32_t* order);
int gg_set_tuning(gg_ctx* ctx, const int32_t* variants);

/* Wait for the stream and report device-side capacity errors. */
int gg_synchronize(gg_ctx* ctx);

/* Grain-count trajectory recorded by each timestep's count stage: writes up to
 * cap entries, *n receives the number of recorded steps. */
int gg_get_count_history(gg_ctx* ctx, int32_t* counts, int32_t cap, int32_t* n);

/* ---- inspection (copies device state to the host; synchronising) ---- */

/* Per-grain state (enum gg_grain_state), grain volume and surface area
 * (LSbox::m_volume / m_surface, src/grainHull.cpp:40-56, 120-131); arrays of
 * n_grains+1.  Any pointer may be NULL. */
int gg_get_grain_stats(gg_ctx* ctx, int32_t* state, double* volume, double* area);

/* Bounds of grain id's input (which=0) or output (which=1) buffer. */
int gg_get_bounds(gg_ctx* ctx, int32_t id, int32_t which, int32_t* bounds6);

/* Copy grain id's distance buffer (which: 0 input, 1 output), converted to
 * double, into dst (capacity cap elements); *n receives the element count. */
int gg_download_distance(gg_ctx* ctx, int32_t id, int32_t which, double* dst, int64_t cap, int64_t* n);

/* Overwrite grain id's distance buffer (which: 0 input, 1 output) with the n
 * doubles at src, x fastest, n = the buffer's box volume (rounded to float in
 * the FP32 path); replaces writing an LSbox's DimensionalBufferReal through
 * setValueAt (src/dimensionalBuffer.h:87-92), e.g. as
 * LSbox::calculateDistanceFunction does (src/box.cpp:374-397).  Synchronises
 * the context stream first; the grain must be alive. */
int gg_upload_phi(gg_ctx* ctx, int32_t id, int32_t which, const double* src, int64_t n);

/* Copy grain id's IDLocal buffer (LSbox::m_IDLocal) and its bounds. */
int gg_download_idlocal(gg_ctx* ctx, int32_t id, int32_t* bounds6, uint32_t* dst, int64_t cap, int64_t* n);

/* Copy grain id's m_secondOrderNeighbours (which=0) or m_comparisonList (which=1). */
int gg_download_list(gg_ctx* ctx, int32_t id, int32_t which, int32_t* dst, int32_t cap, int32_t* n);

/* Bulk forms of gg_get_bounds and gg_download_list for inspection (no
 * reference counterpart: the reference reads these fields per LSbox).  Bounds:
 * rows 0..n_grains of 6.  Lists: CSR with offsets[0..n_grains+1] (row g =
 * grain g, row 0 empty), ids filled up to cap entries, *n = total entries;
 * ids may be NULL to query *n. */
int gg_get_all_bounds(gg_ctx* ctx, int32_t which, int32_t* bounds);
int gg_download_all_lists(gg_ctx* ctx, int32_t which, int32_t* offsets, int32_t* ids, int64_t cap, int64_t* n);

/* Hull topology: the grains each grain's marching-cubes hull touches
 * (MarchingCubesAlgorithm::generateAdditionalInformation / recordNewGrains,
 * src/marchingCubes.cpp:501-602, with LSbox::getNeighbourAt, src/box.cpp:3000-3006):
 * every cut cell names the grain itself at its inside corners and the IDLocal id
 * at the others.  Off by default; while on, every level set (gg_level_set,
 * gg_timestep) collects it, at most 64 grains per hull (more sets a device
 * error reported at the next synchronisation).  Enabling drops a captured
 * timestep graph. */
int gg_enable_hull_topology(gg_ctx* ctx, int32_t enable);
/* Per grain after the last level set: the number of grains its hull touched,
 * itself included (GrainHull::getAllNeighborsCount, which LSbox::extractContour
 * stores in m_neighborCount, src/box.cpp:1154) and whether the domain wall
 * (id 0) is among them (GrainHull::IsNeighbor(0)).  Arrays of n_grains+1
 * entries; 0 / 0 for grains whose hull has no cut cell. */
int gg_get_hull_neighbour_counts(gg_ctx* ctx, int32_t* counts, uint8_t* touches_wall);
/* The grains the hull of grain id touched, in the order the reference records
 * them (first appearance in the (z, y, x) cell loop, ascending within a
 * cell): up to cap ids into dst, *n = their number. */
int gg_download_hull_neighbours(gg_ctx* ctx, int32_t id, int32_t* dst, int32_t cap, int32_t* n);
/* Per grain after the last level set: the distinct per-cell id lists of its
 * hull (MarchingCubesAlgorithm::insertNeighborList, src/marchingCubes.cpp:604-617)
 * counted as GrainHull::computeGrainBoundaryElements classifies them
 * (src/grainHull.cpp:187-226): counts4[4g+0] lists of 2 grains (grain
 * boundaries), +1 of 3 (triple lines), +2 of 4 (quadruple junctions), +3 any
 * other size (higher-order junctions); before GrainHull::mergeJunction.  At
 * most 512 lists per hull (more sets the topology device error). */
int gg_get_hull_junction_counts(gg_ctx* ctx, int32_t* counts4);
/* GrainHull::get_Faces (src/grainHull.cpp:1330-1338) of every alive grain, as
 * grainhdl::save_Texture writes them to Faces_<loop>.bin
 * (src/grainhdl.cpp:762-771, struct Face src/Structs.h:12-19): one face per
 * grain-boundary list (the grain and one neighbour) of the grain's last hull
 * whose neighbour id is smaller than the grain's, with the boundary's area
 * (GrainBoundary::computeArea, src/InterfacialElement.cpp:81-92: the summed
 * triangle areas times h^2).  Grains ascending, a grain's faces by neighbour
 * id (the reference lists them in its elements' order).  The areas are summed
 * with atomics, so their last bits vary from run to run.  Needs hull
 * topology on.  grain_a/grain_b/area: cap entries each, or NULL to count;
 * *n receives the number of faces. */
int gg_get_faces(gg_ctx* ctx, int32_t* grain_a, int32_t* grain_b, double* area, int64_t cap, int64_t* n);
/* The interfacial elements' census of every grain after mergeJunction, what
 * LSbox::get_NumberOfTriplelines / get_NumberOfQuadruplePoints /
 * get_NumberOfHighOrderJunctions report (src/box.h, src/grainHull.cpp:187-274,
 * 483-531): counts4[4*id .. 4*id+3] = grain boundaries, triple lines, quadruple
 * junctions, high-order junctions of the grain's last hull (0 for grains not
 * alive).  Built with the convolution correction on (gg_set_physics);
 * gg_get_hull_junction_counts gives the lists before merging. */
int gg_get_element_counts(gg_ctx* ctx, int32_t* counts4);

/* New bounds from executeSetComparison (m_new{X,Y,Z}{Min,Max}), rows of 6. */
int gg_get_new_bounds(gg_ctx* ctx, int32_t* bounds);

/* Sum of n^3 over the boxes of alive grains (the voxel-updates of one step). */
int gg_box_voxels(gg_ctx* ctx, int64_t* voxels);

/* HIP-event timing on the context stream: with timing enabled every kernel
 * segment below is bracketed by an event pair (and gg_timestep runs eagerly,
 * without the graph).  Segments:
 *  0 convolution (x,y) forward DFT | x,y circulant   1 convolution z
 *  2 convolution (y,x) inverse DFT (FP64)            3 IDLocal resize
 *  4 comparison lists      5 comparison, boxes n <= 32      6 hull reductions
 *  7 level-set finalize    8 work plan / layout      9 redistancing x/y sweeps
 * 10 redistancing z sweeps + clamp   11 input-buffer resize   12 grain count
 * 13 comparison, 32 < n <= 48   14 comparison, 48 < n <= 64
 * 15 comparison, 64 < n <= 96   16 grid-coarsening resampling
 * 17 voxel container (clear + scatter)
 * 18 convolution correction   19 interfacial elements of the hull */
int gg_enable_stage_timing(gg_ctx* ctx, int32_t enable);
#define GG_NUM_SEGMENTS 20
/* Accumulated milliseconds and number of timed passes per segment since the
 * last reset; fills up to n entries (launches may be NULL). */
int gg_get_stage_times(gg_ctx* ctx, double* ms, int64_t* launches, int32_t n, int32_t reset);

#ifdef __cplusplus
}
#endif

#endif /* GG_LEVELSET_H */
