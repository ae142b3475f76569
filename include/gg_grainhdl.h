/*
 * gg_grainhdl.h -- host-side set-up that feeds gg_levelset.h: grid/time
 * variables, the seeded Voronoi microstructure and the initial neighbour
 * lists.  Pure host code (no GPU needed); part of the same shared library.
 *
 * Paths relative to the GraGLeS_3D source tree.
 */
#ifndef GG_GRAINHDL_H
#define GG_GRAINHDL_H

#include <stdint.h>
#include "gg_levelset.h"

#ifdef __cplusplus
extern "C" {
#endif

/* grainhdl::initializeSimulation (src/grainhdl.cpp:68-73) +
 * updateGridAndTimeVariables in Gaussian mode (src/grainhdl.cpp:1419-1444) +
 * the Voronoi-mode decrement of realDomainSize (src/grainhdl.cpp:237-239):
 *   rds = int(N^(1/3) * (4pi/3)^(1/3) / 2 * points_per_grain + 1)
 *   dt = time_step_size / rds^2, h = 1/rds, delta = border/rds,
 *   ngridpoints = rds + 2*border, conv_domain_size = rds - 1.
 * Fills every field of *out except precision/max_list/device/arena_voxels,
 * which are set to GG_FP64_PLAIN/0/0/0. */
int gg_derive_params(int32_t n_grains, int32_t points_per_grain, int32_t border, double time_step_size,
                     gg_params* out);

/* resizeToCube (src/dimensionalBuffer.h:139-219) applied to one bounds row. */
void gg_resize_to_cube(int32_t* bounds6, int32_t ngridpoints);

/* Seeded Voronoi microstructure; replaces VOROMicrostructure
 * (src/grainhdl.cpp:226-396) and the bounding-box part of the LSbox
 * constructor (src/box.cpp:185-211).
 *  - seeds: n_grains points uniform in [0,1)^3 from splitmix64(seed), drawn
 *    x, y, z per grain; grain id = seed index + 1;
 *  - ID field (ngridpoints^3, x fastest): 0 on the border shell (i, j or k
 *    < grid_blowup or >= ngridpoints-1-grid_blowup, src/grainhdl.cpp:366-370),
 *    else the id of the nearest seed to ((j-gb)h, (i-gb)h, (k-gb)h), ties to
 *    the lowest id (non-periodic container, as the reference's);
 *  - raw bounds: [min - gb, max + 1 + gb) of each grain's labelled voxels,
 *    clipped to [0, ngridpoints]; exists = grain has a labelled voxel.
 * idfield may be NULL (bounds only are produced then at the same cost). */
int gg_generate_voronoi(const gg_params* params, uint64_t seed, int32_t* idfield, int32_t* bounds,
                        uint8_t* exists, int32_t nthreads);

/* The n_grains seed points gg_generate_voronoi uses for `seed`, as
 * xyz[3*i .. 3*i+2] for grain id i+1. */
int gg_voronoi_seeds(int32_t n_grains, uint64_t seed, double* xyz);

/* Initial neighbour lists, as grainhdl::find_neighbors builds them with an
 * R-tree over the box bounds (src/grainhdl.cpp:1029-1054,
 * src/box.cpp:2871-2887): for every existing grain, the other existing grains
 * whose cubed boxes overlap it with inclusive bounds.  Order is ascending id
 * (the R-tree's traversal order is not reproduced).  CSR output: offsets has
 * n_grains+2 entries; ids has capacity cap; *n receives the total.  Returns
 * GG_ERR_CAPACITY if cap is too small (with *n set). */
int gg_initial_neighbour_lists(const gg_params* params, const int32_t* bounds, const uint8_t* exists,
                               int32_t* offsets, int32_t* ids, int64_t cap, int64_t* n);

#ifdef __cplusplus
}
#endif

#endif /* GG_GRAINHDL_H */
