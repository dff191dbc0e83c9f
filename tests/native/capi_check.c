/* A compiled C caller of librxhip.so through include/rxhip.h only, as a
 * RenderX maintainer's adapter of IntersectionGPU would be: the struct layouts
 * the Python bindings and the adapters assume are checked at compile time,
 * then the entry points are called.
 *   capi_check cpu : host-only entry points; rx_create must report
 *                    RX_ERR_NO_DEVICE when no gfx950 device is present
 *   capi_check gpu : rx_create, rx_tree_build / serialize, rx_load_tree,
 *                    rx_intersect with and without the last_tid skip
 * Exits 0 when every check holds. */
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "../../include/rxhip.h"

_Static_assert(sizeof(rx_object_desc) == 64, "rx_object_desc is 64 bytes");
_Static_assert(offsetof(rx_object_desc, color) == 16 && offsetof(rx_object_desc, exponent) == 28 &&
                   offsetof(rx_object_desc, color2) == 32 && offsetof(rx_object_desc, ior) == 44 &&
                   offsetof(rx_object_desc, g) == 48,
               "rx_object_desc field offsets");
_Static_assert(sizeof(rx_camera_desc) == 48 && offsetof(rx_camera_desc, sight_dist) == 36 &&
                   offsetof(rx_camera_desc, height) == 44,
               "rx_camera_desc layout");
_Static_assert(sizeof(rx_render_params) == 32 && offsetof(rx_render_params, seed) == 16 &&
                   offsetof(rx_render_params, flags) == 28,
               "rx_render_params layout");
_Static_assert(sizeof(rx_render_tuning) == 28 && offsetof(rx_render_tuning, stack_depth) == 12 &&
                   offsetof(rx_render_tuning, block_size) == 16 && offsetof(rx_render_tuning, camera_rays) == 20 &&
                   offsetof(rx_render_tuning, tail_bounce) == 24,
               "rx_render_tuning layout");
_Static_assert(sizeof(rx_render_stats) == 624 && offsetof(rx_render_stats, ms_total) == 56 &&
                   offsetof(rx_render_stats, extend_launches) == 88 && offsetof(rx_render_stats, guide_rows) == 96 &&
                   offsetof(rx_render_stats, ms_extend_bounce) == 112 &&
                   offsetof(rx_render_stats, extend_rays_bounce) == 240 &&
                   offsetof(rx_render_stats, box_tests_bounce) == 368 &&
                   offsetof(rx_render_stats, tri_tests_bounce) == 496 && RX_STATS_BOUNCES == 16,
               "rx_render_stats layout");
_Static_assert(sizeof(rx_guide_params) == 40, "rx_guide_params layout");

static int fails = 0;
#define CHECK(cond, what)                                   \
  do {                                                      \
    if (!(cond)) {                                          \
      printf("FAIL %s (line %d)\n", what, __LINE__);       \
      fails++;                                              \
    } else {                                                \
      printf("ok   %s\n", what);                           \
    }                                                       \
  } while (0)

/* A unit quad in the plane z = 1 (two triangles, ids 0 and 1) and a second
 * one at z = 2 (ids 2 and 3). */
static const float verts[] = {0, 0, 1, 1, 0, 1, 1, 1, 1, 0, 1, 1, 0, 0, 2, 1, 0, 2, 1, 1, 2, 0, 1, 2};
static const uint32_t tris[] = {0, 1, 2, 0, 2, 3, 4, 5, 6, 4, 6, 7};

static int host_checks(rx_tree** out) {
  rx_tree* t = NULL;
  CHECK(rx_tree_build(8, verts, 4, tris, RX_TREE_ABS_DOUBLE, &t) == RX_OK && t, "rx_tree_build");
  if (!t) return 1;
  uint32_t nn = 0, nl = 0, depth = 0;
  CHECK(rx_tree_info(t, &nn, &nl, &depth) == RX_OK && nl >= 4 && nn >= 1 && nn <= 16, "rx_tree_info");
  printf("     %u nodes, %u leaf triangles, depth %u\n", nn, nl, depth);
  CHECK(rx_tree_build(8, verts, 4, tris, 0, NULL) == RX_ERR_INVALID, "rx_tree_build rejects a NULL output");
  const uint32_t bad[] = {0, 1, 9};
  rx_tree* tb = NULL;
  CHECK(rx_tree_build(8, verts, 1, bad, 0, &tb) == RX_ERR_INVALID, "rx_tree_build rejects a vertex index out of range");
  double rmse = -1;
  const float a[3] = {1, 2, 3}, b[3] = {1, 2, 5};
  /* one pixel: sqrt(|ref - img|^2) over the RGB difference (cmp/compare.cpp:85-124) */
  CHECK(rx_image_rmse(1, 1, a, b, NULL, &rmse) == RX_OK && rmse == 2.0, "rx_image_rmse");
  *out = t;
  return 0;
}

static int gpu_checks(rx_tree* t) {
  rx_ctx* ctx = NULL;
  int rc = rx_create(0, &ctx);
  CHECK(rc == RX_OK && ctx, "rx_create");
  if (rc != RX_OK) return 1;
  uint32_t nn = 0, nl = 0;
  rx_tree_info(t, &nn, &nl, NULL);
  float nodes4[4 * 16], min4[4 * 16], max4[4 * 16], v1[16], v2[16], v3[16];
  if (nn > 16) return 1;
  CHECK(rx_tree_serialize(t, nodes4, min4, max4, v1, v2, v3) == RX_OK, "rx_tree_serialize");
  CHECK(rx_load_tree(ctx, nn, nodes4, min4, max4, nl, v1, v2, v3) == RX_OK, "rx_load_tree");
  /* rays straight up through both quads, from below and from on the first quad */
  const float org[6] = {0.25f, 0.7f, 0.0f, 0.6f, 0.3f, 1.0f};
  const float dir[6] = {0, 0, 1, 0, 0, 1};
  float dist[2] = {0, 0};
  int32_t tid[2] = {0, 0};
  CHECK(rx_intersect(ctx, 2, org, dir, NULL, NULL, NULL, RX_ISECT_ORDERED, dist, tid, NULL, NULL) == RX_OK,
        "rx_intersect");
  CHECK(tid[0] == 1 && dist[0] == 1.0f, "ray from below hits the near quad's triangle 1 at distance 1");
  CHECK((tid[1] == 0 || tid[1] == 1) && dist[1] == 0.0f, "ray on the near quad hits it at distance 0");
  const int32_t skip[2] = {-1, tid[1]};
  CHECK(rx_intersect(ctx, 2, org, dir, NULL, NULL, skip, RX_ISECT_ORDERED, dist, tid, NULL, NULL) == RX_OK,
        "rx_intersect with last_tid");
  CHECK(tid[0] == 1 && tid[1] >= 2 && dist[1] == 1.0f, "last_tid skips the start triangle: the far quad at 1");
  CHECK(rx_intersect(ctx, 2, org, dir, NULL, NULL, NULL, RX_ISECT_LEGALITY, dist, tid, NULL, NULL) == RX_ERR_INVALID,
        "legality without inside/contact is rejected");
  printf("     last error: %s\n", rx_last_error(ctx));
  CHECK(rx_synchronize(ctx) == RX_OK, "rx_synchronize");
  rx_destroy(ctx);
  return 0;
}

int main(int argc, char** argv) {
  const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
  rx_tree* t = NULL;
  host_checks(&t);
  if (gpu) {
    if (t) gpu_checks(t);
  } else {
    rx_ctx* ctx = NULL;
    const int rc = rx_create(0, &ctx);
    CHECK(rc == RX_ERR_NO_DEVICE && ctx == NULL, "rx_create without a gfx950 device reports RX_ERR_NO_DEVICE");
  }
  rx_tree_free(t);
  printf("%s\n", fails ? "FAILED" : "ALL OK");
  return fails ? 1 : 0;
}
