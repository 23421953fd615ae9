/* The C ABI used from C99 (no C++ and no Python): the header compiles as C, the library links, and
 * the host-side entry points behave as documented without a GPU (construction, readback, error
 * codes, ownership).  Prints "ok" or the first failed check. */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "feriphys_gpu.h"

#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) {                                                       \
      printf("failed: %s (line %d): %s\n", #c, __LINE__, fp_last_error()); \
      return 1;                                                       \
    }                                                                 \
  } while (0)

int main(void) {
  /* Cloth::new(32, 32, 0.05, origin, 0.01, ...), top corners pinned (springy/cloth.rs:20-113) */
  const float origin[3] = {0.0f, 0.0f, 0.0f};
  const uint64_t pins[2] = {992, 1023};
  fp_mesh* cloth = NULL;
  CHECK(fp_cloth_new(32, 32, 0.05f, origin, 0.01f, 25.0f, 0.1f, 21.2f, 0.0707f, 5.0f, 0.05f, pins, 2, &cloth) ==
        FP_OK);
  uint64_t np = 0, ns = 0, nf = 0, npin = 0;
  CHECK(fp_mesh_counts(cloth, &np, &ns, &nf, &npin) == FP_OK);
  CHECK(np == 1024 && ns == 4745 && nf == 1922 && npin == 2);
  static float rec[1024 * 10];
  CHECK(fp_mesh_get_points(cloth, rec) == FP_OK);
  CHECK(rec[0] == 0.01f);                        /* Point::mass */
  CHECK(rec[10 * 33 + 1] == 1.0f * 0.05f - 0.8f); /* vertex (1, 1): x = (c * s - cols / 2 * s) + 0 */

  /* errors: NULL output, index out of range, wrong record count; the message is per thread */
  CHECK(fp_cloth_new(2, 2, 1.0f, origin, 1.0f, 1, 1, 1, 1, 1, 1, NULL, 0, NULL) == FP_EINVAL);
  CHECK(strlen(fp_last_error()) > 0);
  CHECK(fp_mesh_add_strut(cloth, 0, 5000, 1.0f, 1.0f) == FP_EINVAL);
  CHECK(fp_mesh_set_points(cloth, rec, 1000) == FP_EINVAL);

  /* SpringyMesh::new on one triangle: mass 1/3 per vertex, L = 2 for a (0,0,0)-(2,0,0) edge */
  const float tri[9] = {0, 0, 0, 2, 0, 0, 0, 2, 0};
  const uint64_t idx[3] = {0, 1, 2};
  fp_mesh* m = NULL;
  CHECK(fp_mesh_new(tri, 3, idx, 3, 1.0f, 1.0f, 1.5f, NULL, NULL, 0, &m) == FP_OK);
  float kcl[9];
  CHECK(fp_mesh_get_struts(m, NULL, kcl, NULL) == FP_OK);
  CHECK(kcl[2] == 2.0f && kcl[0] == 0.5f && kcl[1] == 0.75f);

  /* the simulation takes ownership of its meshes on every path; without a device it reports it */
  fp_obstacle* floor_ = NULL;
  const float fp_[12] = {-1, 0, -1, 1, 0, -1, 1, 0, 1, -1, 0, 1};
  const uint64_t fi[6] = {0, 1, 2, 0, 2, 3};
  CHECK(fp_obstacle_new(fp_, 4, fi, 6, &floor_) == FP_OK);
  fp_mesh* meshes[2] = {cloth, m};
  fp_sim* sim = NULL;
  const int rc = fp_sim_new(meshes, 2, &floor_, 1, &sim);
  CHECK(rc == FP_OK || rc == FP_ENODEV);
  if (rc == FP_OK) {
    float ts = 0.0f;
    CHECK(fp_sim_step(sim, 1, &ts) == FP_OK && ts > 0.0f);
    fp_sim_free(sim);
  } else {
    CHECK(sim == NULL);
  }
  printf("ok\n");
  return 0;
}
