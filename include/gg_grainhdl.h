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

/* Seeds and radii of an uneven (bimodal) grain population, for the power
 * diagram below; no counterpart in the reference, whose VOROMicrostructure
 * places uniform seeds for voro++ (src/grainhdl.cpp:226-396).  Per grain
 * (id i+1, splitmix64(seed)): large with probability frac_large, radius =
 * (size_ratio if large else 1) * exp(sigma * z), z standard normal (Box-Muller);
 * radii scaled so the balls' volumes add up to the unit cube's.  Seeds are
 * placed largest first by random sequential addition: a seed is kept at
 * distance >= a_i r_i + a_j r_j from every earlier seed j, a_i = 0.85 (shrunk
 * by 0.9 after each 200 failed draws for that grain).  xyz: 3 per grain;
 * radius: a_i r_i, so the balls (xyz, radius) are disjoint and each lies
 * inside its own cell of the power diagram of these weights (no cell is empty). */
int gg_bimodal_seeds(int32_t n_grains, uint64_t seed, double frac_large, double size_ratio, double sigma,
                     double* xyz, double* radius);

/* Power (Laguerre) diagram microstructure: each interior voxel belongs to the
 * grain minimising |x - s_i|^2 - radius_i^2 (ties to the lowest id); border
 * shell, voxel coordinates, ID field and bounds exactly as
 * gg_generate_voronoi, to which it reduces (bit for bit) when every radius is
 * 0.  Exact labels by blocks of 8^3 voxels: only the seeds whose smallest
 * power distance over the block can beat every seed's largest are tested. */
int gg_generate_power(const gg_params* params, const double* xyz, const double* radius, int32_t* idfield,
                      int32_t* bounds, uint8_t* exists, int32_t nthreads);

/* Initial neighbour lists, as grainhdl::find_neighbors builds them with an
 * R-tree over the box bounds (src/grainhdl.cpp:1029-1054,
 * src/box.cpp:2871-2887): for every existing grain, the other existing grains
 * whose cubed boxes overlap it with inclusive bounds.  Order is ascending id
 * (the R-tree's traversal order is not reproduced).  CSR output: offsets has
 * n_grains+2 entries; ids has capacity cap; *n receives the total.  Returns
 * GG_ERR_CAPACITY if cap is too small (with *n set). */
int gg_initial_neighbour_lists(const gg_params* params, const int32_t* bounds, const uint8_t* exists,
                               int32_t* offsets, int32_t* ids, int64_t cap, int64_t* n);

/* ---- grain-boundary model of the convolution correction (gg_set_physics) ---- */

/* Misorientation of two cubic orientations given as unit quaternions
 * (q0, q1, q2, q3): myQuaternion::misorientationCubicQxQ
 * (src/myQuaternion.cpp:309-367) of qa^-1 qb, floored at 1 degree as
 * LSbox::computeMisorientation does (src/box.cpp:2704-2719); radians. */
double gg_misorientation(const double* qa, const double* qb);

/* Grain-boundary energy and mobility of every ordered pair (a, b) of ids
 * 0..n_ids-1, at gam[a*n_ids+b] and mob[a*n_ids+b]: Read-Shockley with
 * theta_ref = 15 degrees (1 if unique_energies) and
 * 1 - 0.99 exp(-5 (theta/15deg)^9) floored at 0.1 (1 without the mobility
 * model), of the misorientation above (InterfacialElement::
 * computeReadShockleyEnergy / computeMobilityMisori,
 * src/InterfacialElement.cpp:23-50).  quat: 4 doubles per id. */
int gg_gb_pair_tables(int32_t n_ids, const double* quat, int32_t mobility_model, int32_t unique_energies,
                      double* gam, double* mob, int32_t nthreads);

/* Bunge (phi1, PHI, phi2) Euler angles to a quaternion
 * (myQuaternion::euler2Quaternion, src/myQuaternion.cpp:168-191). */
void gg_euler_to_quaternion(const double* euler, double* q);

/* Seeded orientations, 4 doubles per id in quat[0 .. 4*(n_grains+1)): id 0 (the
 * domain wall) gets the reference's boundary grain orientation, Euler
 * (0, 35 deg, 0) (src/grainhdl.cpp:86); ids 1..n_grains Shoemake's uniform
 * random unit quaternions from splitmix64(seed).  reference_abs != 0 takes
 * the absolute value of every component before normalising, as the
 * reference's randomOriShoemakeQuat does (src/myQuaternion.cpp:291-307), which
 * confines the orientations to one orthant (not uniform). */
int gg_random_orientations(int32_t n_grains, uint64_t seed, int32_t reference_abs, double* quat);

/* Seeded lognormal stored elastic energies in normalised units (the value
 * LSbox's constructor stores after scaling by DislocEnPerM / HAGB_Energy *
 * Physical_Domain_Size, src/box.cpp:232-233): see[g] = median * exp(sigma *
 * N(0,1)) for g = 1..n_grains, see[0] = 0 (the wall). */
int gg_lognormal_energies(int32_t n_grains, uint64_t seed, double median, double sigma, double* see);

#ifdef __cplusplus
}
#endif

#endif /* GG_GRAINHDL_H */
