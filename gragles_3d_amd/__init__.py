"""MI355X-native per-grain level-set timestep of GraGLeS_3D.

The compute lives in libgg_levelset.so (HIP for gfx950, C ABI in
include/gg_levelset.h); this package only binds it.
"""
from .levelset import (ALIVE, CMP_ORDER_GRAIN, CMP_ORDER_SPACE, CONV_DEFAULT, CONV_SPLIT_BF16, DELETED,  # noqa: F401
                       FP32_BATCHED, FP64_PLAIN,
                       GGError, GGParams, GGPhysics, LevelSetContext, ORDER_SPATIAL, ORDER_VOLUME, VANISHED,
                       bimodal_seeds, coarsening_plan, derive_params, euler_to_quaternion, gb_pair_tables,
                       generate_power, generate_voronoi, initial_neighbour_lists, lib, lognormal_energies,
                       misorientation, random_orientations, resize_to_cube, setup_simulation, textured_orientations,
                       uds_text, voronoi_seeds)
