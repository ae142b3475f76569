/*
 * A C program that drives the level-set library through include/gg_levelset.h
 * and include/gg_grainhdl.h only, the way grainhdl's run_sim would
 * (src/grainhdl.cpp:836-919): host set-up, create, boxes, initial distance,
 * neighbour lists, timesteps, statistics, destroy.  It compiles against the
 * headers with a C compiler, so a layout drift between the headers and the
 * library (or the Python binding) shows up here.
 *
 *   abi_check <n_grains> <points_per_grain> <border> <seed> <steps> <precision>
 *
 * Prints one "key value" line per result (doubles with %.17g).  Exit status 0
 * on success, 77 when no GPU is usable (the host set-up was still checked),
 * 1 on any other failure.
 */
#include <stdio.h>
#include <stdlib.h>

#include "gg_grainhdl.h"
#include "gg_levelset.h"

static int check(int rc, const char* what, gg_ctx* ctx) {
  if (rc != GG_OK) {
    fprintf(stderr, "%s failed: %d %s\n", what, rc, ctx ? gg_last_error(ctx) : "");
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc != 7) {
    fprintf(stderr, "usage: %s n_grains points_per_grain border seed steps precision\n", argv[0]);
    return 1;
  }
  const int n = atoi(argv[1]), ppg = atoi(argv[2]), border = atoi(argv[3]);
  const unsigned long long seed = strtoull(argv[4], NULL, 10);
  const int steps = atoi(argv[5]), precision = atoi(argv[6]);

  gg_params p;
  if (check(gg_derive_params(n, ppg, border, 0.2, &p), "gg_derive_params", NULL)) return 1;
  p.precision = precision;
  printf("ngridpoints %d\nh %.17g\ndelta %.17g\ndt %.17g\n", p.ngridpoints, p.h, p.delta, p.dt);

  const size_t L = (size_t)p.ngridpoints;
  int32_t* idf = malloc(sizeof(int32_t) * L * L * L);
  int32_t* bounds = calloc((size_t)6 * (n + 1), sizeof(int32_t));
  uint8_t* exists = calloc((size_t)n + 1, 1);
  int32_t* offs = calloc((size_t)n + 2, sizeof(int32_t));
  if (!idf || !bounds || !exists || !offs) return 1;
  if (check(gg_generate_voronoi(&p, seed, idf, bounds, exists, 0), "gg_generate_voronoi", NULL)) return 1;
  int64_t nl = 0;
  gg_initial_neighbour_lists(&p, bounds, exists, offs, NULL, 0, &nl); /* query the total */
  int32_t* ids = malloc(sizeof(int32_t) * (size_t)(nl > 0 ? nl : 1));
  if (check(gg_initial_neighbour_lists(&p, bounds, exists, offs, ids, nl, &nl), "gg_initial_neighbour_lists", NULL))
    return 1;
  long long idsum = 0;
  for (size_t v = 0; v < L * L * L; ++v) idsum += idf[v];
  printf("idfield_sum %lld\nlist_entries %lld\n", idsum, (long long)nl);

  gg_ctx* ctx = NULL;
  int rc = gg_create(&p, &ctx);
  if (rc == GG_ERR_DEVICE) {
    printf("device none\n");
    return 77;
  }
  if (check(rc, "gg_create", NULL)) return 1;
  if (check(gg_set_boxes(ctx, bounds, exists), "gg_set_boxes", ctx) ||
      check(gg_init_distance(ctx, idf), "gg_init_distance", ctx) ||
      check(gg_set_neighbour_lists(ctx, offs, ids), "gg_set_neighbour_lists", ctx) ||
      check(gg_timestep(ctx, steps, 1), "gg_timestep", ctx) || check(gg_synchronize(ctx), "gg_synchronize", ctx))
    return 1;

  int32_t hist[4096];
  int32_t nh = 0;
  if (check(gg_get_count_history(ctx, hist, 4096, &nh), "gg_get_count_history", ctx)) return 1;
  printf("count_history");
  for (int i = 0; i < nh && i < 4096; ++i) printf(" %d", hist[i]);
  printf("\n");

  int32_t* state = calloc((size_t)n + 1, sizeof(int32_t));
  double* vol = calloc((size_t)n + 1, sizeof(double));
  double* area = calloc((size_t)n + 1, sizeof(double));
  if (check(gg_get_grain_stats(ctx, state, vol, area), "gg_get_grain_stats", ctx)) return 1;
  for (int g = 1; g <= n; ++g) printf("grain %d %d %.17g %.17g\n", g, state[g], vol[g], area[g]);

  /* the output buffer of the first alive grain: bounds and the sum of its values */
  for (int g = 1; g <= n; ++g) {
    if (state[g] != GG_ALIVE) continue;
    int32_t b[6];
    int64_t cnt = 0;
    if (check(gg_get_bounds(ctx, g, 1, b), "gg_get_bounds", ctx) ||
        check(gg_download_distance(ctx, g, 1, NULL, 0, &cnt), "gg_download_distance", ctx))
      return 1;
    double* phi = malloc(sizeof(double) * (size_t)cnt);
    if (check(gg_download_distance(ctx, g, 1, phi, cnt, &cnt), "gg_download_distance", ctx)) return 1;
    double s = 0.0;
    for (int64_t v = 0; v < cnt; ++v) s += phi[v];
    printf("phi_grain %d %d %d %d %d %d %d %lld %.17g\n", g, b[0], b[1], b[2], b[3], b[4], b[5], (long long)cnt, s);
    free(phi);
    break;
  }

  /* a second redistancing without a level set before it is refused */
  rc = gg_redistancing(ctx);
  printf("redistancing_again %d\n", rc);
  if (rc != GG_ERR_STATE) return 1;

  if (check(gg_destroy(ctx), "gg_destroy", NULL)) return 1;
  free(idf);
  free(bounds);
  free(exists);
  free(offs);
  free(ids);
  free(state);
  free(vol);
  free(area);
  return 0;
}
