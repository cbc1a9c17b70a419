/* Drives the CPU oracle (oracle/meso_oracle.c) through its entry points on
 * small inputs, built with AddressSanitizer + UndefinedBehaviorSanitizer by
 * tests/test_native_sanitizers.py: generators, the three scene kinds with
 * primary, shadow and AO rays (one thread and several), edits, scene chunks,
 * the culled and greedy meshers (checkerboard worst case included), the
 * chunk pool over a camera path, quad visibility and the reference
 * restatements.  It checks a few invariants; the sanitizers check the
 * memory and arithmetic. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "meso_oracle.h"

static int failures = 0;
#define EXPECT(c)                                                          \
  do {                                                                     \
    if (!(c)) {                                                            \
      fprintf(stderr, "oracle_sanitize: %s failed at line %d\n", #c, __LINE__); \
      failures++;                                                          \
    }                                                                      \
  } while (0)

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s ? s : 1);
  if (!p) abort();
  return p;
}

static void render(const mo_scene *s, int32_t w, int32_t h, uint32_t flags, int threads) {
  mo_camera cam;
  const float n = (float)s->n;
  const float eye[3] = {-0.4f * n, 1.3f * n, -0.2f * n}, target[3] = {0.5f * n, 0.2f * n, 0.5f * n};
  const float up[3] = {0.0f, 1.0f, 0.0f};
  mo_camera_derive(eye, target, up, 60.0f, w, h, &cam);
  mo_shading sh = {{0.4f, 0.8f, 0.45f}, 7u, 4, 16.0f};
  const size_t px = (size_t)w * h;
  int64_t *idx = xcalloc(px, sizeof(int64_t));
  uint8_t *face = xcalloc(px, 1);
  float *t = xcalloc(px, sizeof(float)), *shade = xcalloc(px, sizeof(float));
  mo_counters c;
  memset(&c, 0, sizeof(c));
  if (threads > 1)
    mo_raymarch_mt(s, &cam, flags, &sh, 0, h, threads, idx, face, t, shade, &c);
  else
    mo_raymarch(s, &cam, flags, &sh, 0, h, idx, face, t, shade, &c);
  EXPECT(c.rays == (int64_t)px);
  for (size_t i = 0; i < px; i++) EXPECT((idx[i] < 0) == (face[i] == MO_FACE_MISS));
  /* an eye inside the scene and a ray along an axis */
  mo_hit hit;
  const float o[3] = {0.5f * n + 0.25f, 0.5f * n, 0.5f * n}, d[3] = {0.0f, -1.0f, 0.0f};
  mo_trace(s, o, d, &hit, &c);
  free(idx);
  free(face);
  free(t);
  free(shade);
}

static void brick_scenes(void) {
  const int32_t n = 128, nb = n / 8;
  uint32_t *slots = xcalloc((size_t)nb * nb * nb, 4);
  const int64_t cnt = mo_build_brickmap(MO_GEN_TERRAIN_CAVES, n, 9u, slots, NULL, NULL, 0);
  EXPECT(cnt > 0);
  uint64_t *masks = xcalloc((size_t)cnt * 8, 8);
  uint16_t *vox = xcalloc((size_t)cnt * 512, 2);
  EXPECT(mo_build_brickmap(MO_GEN_TERRAIN_CAVES, n, 9u, slots, masks, vox, cnt) == cnt);
  const int32_t n32 = (n + 31) / 32, n128 = (n + 127) / 128;
  uint32_t *occ32 = xcalloc(((size_t)n32 * n32 * n32 + 31) / 32, 4);
  uint32_t *occ128 = xcalloc(((size_t)n128 * n128 * n128 + 31) / 32, 4);
  mo_scene s;
  memset(&s, 0, sizeof(s));
  s.kind = MO_SCENE_BRICKMAP_U16;
  s.n = n;
  s.slots = slots;
  s.masks = masks;
  s.voxels = vox;
  mo_build_coarse(&s, 0, occ32, occ128, NULL);
  s.occ32 = occ32;
  s.occ128 = occ128;
  mo_scene_box(&s, s.box_lo, s.box_hi);
  render(&s, 67, 41, 0, 1);
  render(&s, 67, 41, MO_RAY_SHADOW | MO_RAY_AO, 3);

  /* edits, scene chunks, meshes of the edited chunks */
  const int32_t xyz[9] = {0, 0, 0, 1, 2, 3, 3, 3, 3};
  mo_edit_chunks(&s, vox, masks, xyz, 3, 5u);
  uint16_t *chunk = xcalloc(32768, 2);
  uint32_t *quads = xcalloc(98304, 4);
  for (int k = 0; k < 3; k++) {
    mo_scene_chunk(&s, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], chunk);
    const uint32_t nq = mo_mesh_chunk(chunk, 2, 1, quads, 98304);
    const uint32_t nc = mo_mesh_chunk(chunk, 2, 0, quads, 98304);
    EXPECT(nq <= nc);
  }

  /* the chunk pool over a short camera path, then quad visibility */
  mo_pool_params p = {4, 2, mo_pool_view_threshold(100.0f), 40, 2000, 10, 1};
  const int32_t nc = n / 32;
  uint32_t *lookup = xcalloc((size_t)nc * nc * nc, 4);
  for (int32_t i = 0; i < nc * nc * nc; i++) lookup[i] = MO_NO_SLOT;
  uint32_t *slot_chunk = xcalloc((size_t)p.capacity, 4), *slot_stamp = xcalloc((size_t)p.capacity, 4);
  uint32_t *slot_count = xcalloc((size_t)p.capacity, 4);
  for (int32_t i = 0; i < p.capacity; i++) slot_chunk[i] = MO_NO_SLOT;
  uint32_t *pq = xcalloc((size_t)p.capacity * p.quads_per_slot, 4);
  int32_t *load_xyz = xcalloc((size_t)p.max_loads * 3, 4);
  uint32_t *load_slot = xcalloc((size_t)p.max_loads, 4);
  float dirs[3 * 16];
  mo_pool_bake_directions(16, dirs);
  const int32_t path[4][3] = {{1, 1, 1}, {2, 1, 1}, {2, 2, 3}, {-1, 0, 0}};
  for (uint32_t step = 0; step < 4; step++) {
    const float raw[3] = {1.0f, 0.1f * (float)step, -0.3f};
    float f[3], snapped[3];
    mo_pool_forward(raw, f);
    mo_pool_snap(dirs, 16, f, snapped);
    mo_pool_result r;
    mo_pool_update(&s, &p, step + 1, path[step], snapped, lookup, slot_chunk, slot_stamp,
                   slot_count, pq, load_xyz, load_slot, &r);
    EXPECT(r.n_loads <= (uint32_t)p.max_loads);
  }
  mo_camera cam;
  const float eye[3] = {40.0f, 50.0f, 40.0f}, target[3] = {90.0f, 20.0f, 100.0f};
  const float up[3] = {0.0f, 1.0f, 0.0f};
  mo_camera_derive(eye, target, up, 70.0f, 48, 32, &cam);
  const size_t px = 48 * 32;
  uint64_t *keys = xcalloc(px, 8);
  int64_t *idx = xcalloc(px, 8);
  uint8_t *face = xcalloc(px, 1);
  float *t = xcalloc(px, 4), *shade = xcalloc(px, 4);
  const float light[3] = {0.4f, 0.8f, 0.45f};
  mo_raster_pool(&cam, n, nc, slot_chunk, slot_count, pq, p.capacity, p.quads_per_slot, light,
                 keys, idx, face, t, shade);

  /* the same bricks as a two-level scene (one node of 16^3 bricks) */
  uint32_t top[1] = {0};
  uint32_t *node_slots = xcalloc(4096, 4);
  for (int32_t bz = 0; bz < nb; bz++)
    for (int32_t by = 0; by < nb; by++)
      for (int32_t bx = 0; bx < nb; bx++)
        node_slots[bx + 16 * by + 256 * bz] = slots[bx + nb * (by + nb * bz)];
  uint64_t groups[1] = {0};
  mo_scene s2;
  memset(&s2, 0, sizeof(s2));
  s2.kind = MO_SCENE_TWO_LEVEL_U16;
  s2.n = n;
  s2.top = top;
  s2.slots = node_slots;
  s2.masks = masks;
  s2.voxels = vox;
  mo_build_coarse(&s2, 1, NULL, NULL, groups);
  s2.groups = groups;
  mo_scene_box(&s2, s2.box_lo, s2.box_hi);
  render(&s2, 53, 29, MO_RAY_SHADOW | MO_RAY_AO, 1);

  free(keys); free(idx); free(face); free(t); free(shade);
  free(lookup); free(slot_chunk); free(slot_stamp); free(slot_count); free(pq);
  free(load_xyz); free(load_slot); free(node_slots);
  free(chunk); free(quads); free(occ32); free(occ128);
  free(slots); free(masks); free(vox);
}

int main(void) {
  /* dense spheres */
  int32_t params[4 * 8];
  mo_sphere_params(1u, 64, 8, params);
  uint8_t *grid = xcalloc(64 * 64 * 64, 1);
  mo_gen_spheres_dense(params, 8, 64, grid);
  mo_scene d;
  memset(&d, 0, sizeof(d));
  d.kind = MO_SCENE_DENSE_U8;
  d.n = 64;
  d.dense = grid;
  mo_scene_box(&d, d.box_lo, d.box_hi);
  render(&d, 31, 23, MO_RAY_SHADOW, 1);
  render(&d, 31, 23, MO_RAY_SHADOW | MO_RAY_AO, 2);
  free(grid);

  brick_scenes();

  /* meshers on generated chunks and on the checkerboard worst case */
  mo_chunk_params cp[4];
  mo_chunk_params_gen(3u, 4, cp);
  uint16_t *vox = xcalloc(32768, 2);
  uint32_t *quads = xcalloc(98304, 4);
  for (int c = 0; c < 4; c++) {
    mo_gen_chunk(&cp[c], vox);
    EXPECT(mo_mesh_chunk(vox, 2, 1, quads, 98304) <= mo_mesh_chunk(vox, 2, 0, quads, 98304));
  }
  for (int i = 0; i < 32768; i++) vox[i] = (uint16_t)(((i & 31) + ((i >> 5) & 31) + (i >> 10)) & 1);
  EXPECT(mo_mesh_chunk(vox, 2, 0, quads, 98304) == 98304);
  EXPECT(mo_mesh_chunk(vox, 2, 1, quads, 98304) == 98304);
  EXPECT(mo_mesh_chunk(vox, 2, 0, quads, 100) == 98304); /* capacity cuts writes, not the count */
  uint32_t counts[4];
  uint16_t *many = xcalloc(4 * 32768, 2);
  for (int c = 0; c < 4; c++) mo_gen_chunk(&cp[c], many + (size_t)c * 32768);
  mo_mesh_chunks_mt(many, 2, 4, 1, 3, counts);
  free(many);
  free(vox);
  free(quads);

  /* reference restatements */
  uint8_t cells[3 * 4096];
  EXPECT(mo_ref_generate_sphere(3, 0, 0, 1.0f, 16, cells) > 0);
  EXPECT(mo_ref_test_generator(0, -3, 0, 1.0f, 16, cells) == 4096);
  uint8_t occ[4096], mips[4 * 4096];
  for (int i = 0; i < 4096; i++) occ[i] = (uint8_t)((i % 7) != 0);
  mo_ref_erode_mips(occ, 16, 4, mips);
  const double x[3] = {1.5, -2.25, 3.75};
  double nd[4];
  mo_ref_noised(x, nd);
  EXPECT(fabs(nd[3]) <= 1.0);
  EXPECT(fabs(mo_ref_displacement(x)) < 1.0);
  if (failures) {
    fprintf(stderr, "oracle_sanitize: %d failures\n", failures);
    return 1;
  }
  printf("oracle_sanitize ok\n");
  return 0;
}
