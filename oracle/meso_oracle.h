/*
 * meso_oracle.h -- CPU oracle for the MesoEngine voxel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This scalar C restatement is the checker that
 * the HIP library (mesoengine_amd/csrc) is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product never links, loads or calls it.
 *
 * The reference (yuchengzhong/MesoEngine) has no DDA raymarcher and no
 * face-cull/greedy mesher.  The traversal follows Amanatides & Woo (1987),
 * "A Fast Voxel Traversal Algorithm for Ray Tracing".  The greedy merge
 * follows M. Lysenko (2012), "Meshing in a Minecraft game".  Both are
 * restated here over the reference's data conventions:
 *   - linear index x + y*Rx + z*Rx*Ry, x fastest
 *     (Runtimes/Helper/VoxelMathHelper.h:73-76)
 *   - [0,n) bound tests (VoxelMathHelper.h:98-104); outside the grid or
 *     chunk counts as empty (BinaryOccupancyVolume.h:43-51)
 *   - face enumeration 0:+Z 1:-Z 2:+Y 3:-Y 4:+X 5:-X, the vertex-group order
 *     of Runtimes/Shape/Cube.h:25-54
 *   - brick edge 8 (Runtimes/Voxel/VoxelSceneConfig.h:22), a top node of 16^3
 *     bricks (VoxelSceneConfig.h:24), and an 8^3 bit occupancy mask of 64 B
 *     per brick (Runtimes/Voxel/VoxelStructure.h:30-39)
 *   - packed little-endian u32 records (Runtimes/Voxel/Block/Block.h:21-27,
 *     unpacked in Samples/SimpleVoxel.cpp:79-85)
 *
 * Pins (each function below is checked by tests/ against something other
 * than itself):
 *   PCG32 / SplitMix64: published known-answer outputs.
 *   Generators: numpy re-evaluation at integer/fixed-point level (spheres);
 *     closed-form lattice counts (reference sphere); invariants.
 *   DDA: brute-force FP64 slab intersection against every solid voxel on
 *     tiny grids; closed forms (single voxel, axis-aligned box, axis-parallel
 *     rays); FP64 entry-plane t of the chosen voxel.
 *   Shading: FP64 recomputation of every hit pixel's shade (lambert, shadow
 *     and AO occlusion by brute-force slab tests against every solid voxel).
 *   Edits / scene chunks: numpy restatement over the dense grid.
 *   Quad visibility: float64 ray-quad brute force on single and stacked
 *     quads; agreement with the DDA on whole frames of fully resident scenes.
 *   Chunk pool: closed-form importance values; the desired set against
 *     integer lattice counts (equal radii) and cone cases; update invariants
 *     (lookup and slots inverse maps, top-ranked loads, least important
 *     evictions, load limit) against a Python ranking over float64.
 *   Meshers: brute-force face enumeration; closed forms (a x b x c box);
 *     greedy tiling invariants; erosion-cull cross-check.
 *   Reference TestGenerator (sin hash): parity unpinned (CPU timing only).
 *
 * Build: cc -O2 -ffp-contract=off (no -ffast-math), so every float result is
 * the IEEE-754 single-precision result of the operations as written.
 */
#ifndef MESO_ORACLE_H
#define MESO_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- conventions ---------------- */
#define MO_FACE_PZ 0
#define MO_FACE_NZ 1
#define MO_FACE_PY 2
#define MO_FACE_NY 3
#define MO_FACE_PX 4
#define MO_FACE_NX 5
#define MO_FACE_INSIDE 6   /* ray origin inside a solid voxel */
#define MO_FACE_MISS 0xFF
#define MO_EMPTY_SLOT 0xFFFFFFFFu
#define MO_BRICK 8
#define MO_NODE_BRICKS 16  /* top node = 16^3 bricks = 128^3 voxels */
#define MO_CHUNK 32

/* quad = x:5 | y:5 | z:5 | (w-1):5 | (h-1):5 | face:3 | 0:4 */
uint32_t mo_quad_pack(int x, int y, int z, int w, int h, int face);

/* ---------------- RNG and hashing ---------------- */
uint64_t mo_splitmix64(uint64_t *state);
typedef struct { uint64_t state, inc; } mo_pcg32;
void mo_pcg32_srandom(mo_pcg32 *r, uint64_t initstate, uint64_t initseq);
uint32_t mo_pcg32_next(mo_pcg32 *r);
/* seed -> (initstate, initseq) via two SplitMix64 draws */
void mo_pcg32_seed(mo_pcg32 *r, uint64_t seed);
uint32_t mo_hash32(uint32_t x); /* lowbias32 integer finaliser */
uint32_t mo_lattice_hash3(int32_t x, int32_t y, int32_t z, uint32_t seed);
float mo_value_noise2(float x, float z, uint32_t seed);
float mo_value_noise3(float x, float y, float z, uint32_t seed);
/* fBm in [0,1): octaves, lacunarity 2, gain 0.5, normalised by sum of amps */
float mo_fbm2(float x, float z, int octaves, float base_freq, uint32_t seed);
float mo_fbm3(float x, float y, float z, int octaves, float base_freq, uint32_t seed);

/* ---------------- scene generators ---------------- */
/* dense spheres: params[4*i] = cx, cy, cz, r in 1/256 voxel fixed point */
void mo_sphere_params(uint64_t seed, int n, int n_spheres, int32_t *params);
void mo_gen_spheres_dense(const int32_t *params, int n_spheres, int n, uint8_t *grid);

/* heightfield terrain: column height and material rule */
int32_t mo_height(int32_t x, int32_t z, int32_t n, uint32_t seed);
uint16_t mo_heightfield_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed);
/* terrain with carved tunnels (serving scenes) and thin caves (sparse) */
uint16_t mo_terrain_caves_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed);
uint16_t mo_caves_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed);

#define MO_GEN_HEIGHTFIELD 1
#define MO_GEN_TERRAIN_CAVES 2
#define MO_GEN_CAVES 3
uint16_t mo_scene_voxel(int kind, int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed);

/* Brick map of a procedural scene of edge n (multiple of 8).
 * slots[nb^3] (nb = n/8): pool index or MO_EMPTY_SLOT, assigned in linear
 * brick order.  masks[8 u64 per brick]: word z, bit x + 8*y.
 * voxels[512 u16 per brick]: x + 8*y + 64*z.
 * Call with masks == NULL to count only.  Returns the number of bricks. */
int64_t mo_build_brickmap(int kind, int32_t n, uint32_t seed, uint32_t *slots,
                          uint64_t *masks, uint16_t *voxels, int64_t cap_bricks);
/* Brick map from a dense u16 grid (x fastest). */
int64_t mo_brickmap_from_dense(const uint16_t *grid, int32_t n, uint32_t *slots,
                               uint64_t *masks, uint16_t *voxels, int64_t cap_bricks);
/* Fill one brick (bx,by,bz) of a procedural scene; returns solid count. */
int32_t mo_gen_brick(int kind, int32_t n, uint32_t seed, int32_t bx, int32_t by,
                     int32_t bz, uint16_t *vox512, uint64_t *mask8);

/* chunk batch (4096 x 32^3): per-chunk parameters and voxel fill */
typedef struct {
  float fill;           /* Beta(1/2,1/2) draw */
  uint32_t noise_seed;
  int32_t n_spheres;    /* 1..4 */
  int32_t sph[4][4];    /* cx, cy, cz, r in 1/256 voxel fixed point */
} mo_chunk_params;
void mo_chunk_params_gen(uint64_t seed, int32_t n_chunks, mo_chunk_params *out);
void mo_gen_chunk(const mo_chunk_params *p, uint16_t *vox32768);

/* ---------------- camera and rays ---------------- */
typedef struct {
  float eye[3], fwd[3], right[3], up[3];
  float sx, sy, inv_w2, inv_h2;
  int32_t width, height;
} mo_camera;
void mo_camera_derive(const float eye[3], const float target[3], const float up[3],
                      float vfov_deg, int32_t width, int32_t height, mo_camera *cam);
void mo_primary_ray(const mo_camera *cam, int32_t px, int32_t py, float o[3], float d[3]);

/* ---------------- scenes for traversal ---------------- */
#define MO_SCENE_DENSE_U8 1
#define MO_SCENE_BRICKMAP_U16 2
#define MO_SCENE_TWO_LEVEL_U16 3
typedef struct {
  int32_t kind;
  int32_t n;                 /* voxels per axis (cubic scene) */
  const uint8_t *dense;      /* dense u8, n^3 */
  const uint32_t *slots;     /* brick map: (n/8)^3 ; two-level: brick slots of nodes */
  const uint64_t *masks;
  const uint16_t *voxels;
  const uint32_t *top;       /* two-level: (n/128)^3 node indices */
  /* empty-space levels above the bricks (see mo_build_coarse) */
  const uint32_t *occ32;     /* brick map: ceil(n/32)^3 bits */
  const uint32_t *occ128;    /* brick map: ceil(n/128)^3 bits */
  const uint64_t *groups;    /* two-level: per node, bit of each 4^3-brick group */
  /* the box [box_lo, box_hi) that holds every solid voxel (mo_scene_box):
   * rays are clipped to it, walks end when they step out of it, and
   * secondary rays that start outside it are unoccluded */
  int32_t box_lo[3], box_hi[3];
} mo_scene;

/* Empty-space levels.  Brick map: occ32 / occ128 hold one bit per 32^3 /
 * 128^3 voxel block (x fastest, word = bit >> 5), set when any non-empty
 * brick lies in the block.  Two-level: the top node table is the 128 level,
 * and groups[node] bit gx + 4 gy + 16 gz (g = (brick & 15) >> 2) is the 32
 * level.  Arrays: occ32 ceil(ceil(n/32)^3 / 32) words, occ128 likewise,
 * groups n_nodes words. */
void mo_build_coarse(const mo_scene *s, int64_t n_nodes, uint32_t *occ32, uint32_t *occ128,
                     uint64_t *groups);
/* The scene's box: brick map and two-level, the bricks that are allocated,
 * in voxels (lo = 8 * min brick, hi = 8 * (max brick + 1)); empty when no
 * brick is; dense grids, the whole grid. */
void mo_scene_box(const mo_scene *s, int32_t lo[3], int32_t hi[3]);

/* step counters for the algorithmic byte model.  n_vox: voxel reads (dense,
 * 1 byte) or voxel-level 8-byte mask-word reads (brick scenes); n_brickmask:
 * bricks entered (their 8-byte 2^3-block word read) */
typedef struct {
  int64_t n_vox, n_top, n_node, n_brickslot, n_brickmask, hits;
  int64_t rays, shadow_rays, ao_rays;
  int64_t n_coarse;  /* coarse words read: 128^3-block bits (brick map) and
                        each 32^3 block's word of allocated bricks (both) */
  int64_t n_iter;    /* traversal steps at any level */
} mo_counters;

typedef struct {
  int64_t index;   /* -1 on miss */
  int32_t face;    /* 0..5, MO_FACE_INSIDE, MO_FACE_MISS */
  float t;         /* +inf on miss */
  int32_t cell[3];
  uint16_t material;
} mo_hit;

/* First solid voxel along o + t d inside [0,n)^3.  Origin anywhere.  The
 * walk steps at the coarsest empty level that holds the current cell: 128^3
 * and 32^3 blocks, bricks, and inside a non-empty brick its aligned 4^3 and
 * 2^3 blocks, else the voxel.  A step at level L leaves the block through its
 * nearest plane; the other coordinates come from the exit point, clamped into
 * the block and never moving against the ray. */
int mo_trace(const mo_scene *s, const float o[3], const float d[3], mo_hit *hit,
             mo_counters *cnt);
/* Occlusion ray from a point on a face of voxel `cell`, starting in the empty
 * neighbour cell `start`.  Returns 1 when a solid voxel is entered at t < tmax. */
int mo_occluded(const mo_scene *s, const float o[3], const float d[3],
                const int32_t start[3], float tmax, mo_counters *cnt);

#define MO_RAY_SHADOW 1u
#define MO_RAY_AO 2u
typedef struct {
  float light[3];     /* unit vector toward the light */
  uint32_t ao_seed;
  int32_t ao_samples; /* 4 */
  float ao_radius;    /* voxels */
} mo_shading;
/* Renders rows [row0, row1) of the image; out arrays are width*height. */
void mo_raymarch(const mo_scene *s, const mo_camera *cam, uint32_t flags,
                 const mo_shading *sh, int32_t row0, int32_t row1,
                 int64_t *out_index, uint8_t *out_face, float *out_t, float *out_shade,
                 mo_counters *cnt);
/* Same over nthreads POSIX threads (static row partition); counters summed. */
void mo_raymarch_mt(const mo_scene *s, const mo_camera *cam, uint32_t flags,
                    const mo_shading *sh, int32_t row0, int32_t row1, int32_t nthreads,
                    int64_t *out_index, uint8_t *out_face, float *out_t, float *out_shade,
                    mo_counters *cnt);
/* Cosine-hemisphere AO direction (Malley's method, disk by rejection). */
void mo_ao_direction(uint32_t seed, uint32_t pixel, uint32_t sample, int face, float dir[3]);

/* ---------------- meshing ---------------- */
/* chunk of 32^3 voxels, x fastest; elem_size 1 (u8) or 2 (u16); 0 = empty.
 * Quads are emitted in canonical order: face, slice, then within the plane
 * by the quad's last row (v + h - 1) and its start u (the order in which a
 * row sweep closes them; for 1x1 culled quads this is v, u).
 * Returns the quad count (writes at most cap). */
uint32_t mo_mesh_chunk(const void *vox, int elem_size, int greedy, uint32_t *quads,
                       uint32_t cap);
uint64_t mo_mesh_chunks_mt(const void *vox, int elem_size, int32_t n_chunks, int greedy,
                           int32_t nthreads, uint32_t *counts);

/* ---------------- reference restatements ---------------- */
/* FChunk::CalculateOccupancyErodeMipmaps (Runtimes/Voxel/Chunk/Chunk.h:73-94)
 * with FOccupancyHelper::ErodeSingleVoxel (BinaryOccupancyVolume.h:76-99).
 * occ: res^3 bytes (0/1), x fastest; mips: depth * res^3 bytes. */
void mo_ref_erode_mips(const uint8_t *occ, int32_t res, int32_t depth, uint8_t *mips);
/* FGeneratorHelper::GenerateSphere (Runtimes/Helper/GeneratorHelper.h:120-150):
 * solid cells of one chunk in X-outer, Y, Z-inner order; cells[3*k] = x,y,z. */
int32_t mo_ref_generate_sphere(int32_t cx, int32_t cy, int32_t cz, float block_size,
                               int32_t res, uint8_t *cells);
/* FGeneratorHelper::TestGenerator (GeneratorHelper.h:90-119) with noised /
 * displacement (:18-87) and FVoxelMathHelper::Hash (VoxelMathHelper.h:29-33),
 * in double as the reference instantiates them (dvec3).  The hash's bits
 * follow the platform libm sin at large arguments; the structure (lattice
 * values, derivatives, octave composition, the height test) is pinned by
 * tests/test_oracle_refnoise.py. */
int32_t mo_ref_test_generator(int32_t cx, int32_t cy, int32_t cz, float block_size,
                              int32_t res, uint8_t *cells);
/* FVoxelMathHelper::Hash of a dvec3 (VoxelMathHelper.h:29-33) */
double mo_ref_hash3(double x, double y, double z);
/* FGeneratorHelper::noised (GeneratorHelper.h:18-53); out = (dx, dy, dz, value),
 * the .yzwx order the reference returns */
void mo_ref_noised(const double x[3], double out[4]);
/* FGeneratorHelper::displacement (GeneratorHelper.h:55-87) */
double mo_ref_displacement(const double p[3]);
/* GetOctantId (Samples/SimpleVoxel.cpp:129-136) */
int32_t mo_ref_octant_id(float x, float y, float z);

/* ---------------- edits and scene chunks ---------------- */
/* Seeded edit of 32^3 chunks (chunk coordinates xyz[3k]): inside a sphere
 * drawn per chunk from lowbias32(seed ^ lowbias32(chunk linear index)) every
 * voxel of an allocated brick toggles (solid -> 0, empty -> 7); masks follow.
 * Bricks are never allocated or freed.  voxels / masks are modified in place. */
void mo_edit_chunks(const mo_scene *s, uint16_t *voxels, uint64_t *masks, const int32_t *xyz,
                    int32_t n_chunks, uint32_t seed);
/* The 32^3 u16 voxels of chunk (cx, cy, cz) of a brick scene (x fastest). */
void mo_scene_chunk(const mo_scene *s, int32_t cx, int32_t cy, int32_t cz, uint16_t *out);

/* ---------------- chunk pool (streaming around a camera) ---------------- */
/* A fixed number of slots holds the meshed quads of the 32^3 chunks around
 * the camera, chosen by the reference's importance ranking
 * (Runtimes/Voxel/Chunk/ChunkManagerHelper.h:21-43 importance,
 * :91-148 desired set by view, ChunkPool.h:446-560 FindMin replacement,
 * ChunkManager.h:211-260 the per-update load limit with the rest re-queued).
 * Offsets (dx, dy, dz) are chunk offsets from the camera chunk; f is the unit
 * forward vector (mo_pool_forward).  All float expressions are written one
 * IEEE binary32 operation at a time; length = sqrt((x*x + y*y) + z*z) and
 * normalize = v / length. */
#define MO_NO_SLOT 0xFFFFFFFFu
typedef struct {
  int32_t forward_radius;  /* ViewForwardLoadChunkSize, chunks */
  int32_t backward_radius; /* ViewBackwardLoadChunkSize */
  float view_threshold;    /* max(cos(ViewChunkAngle / 2), 0.01) (mo_pool_view_threshold) */
  int32_t capacity;        /* slots */
  int32_t quads_per_slot;  /* quad capacity of a slot; longer meshes are cut (count kept) */
  int32_t max_loads;       /* chunks loaded per update; the rest wait for later updates */
  int32_t greedy;          /* 1 greedy, 0 culled */
} mo_pool_params;
typedef struct {
  uint32_t n_candidates; /* desired, non-empty, not resident */
  uint32_t n_loads;      /* loaded this update */
  uint32_t n_evicted;    /* resident chunks replaced this update */
} mo_pool_result;
float mo_pool_view_threshold(float view_angle_deg);
void mo_pool_forward(const float raw[3], float f[3]);
/* FVoxelMathHelper::GetFibonacciSphere<float> (Runtimes/Helper/VoxelMathHelper.h:47-68),
 * as BakeVisibilityByView instantiates it (ChunkManagerHelper.h:205): n >= 2
 * unit directions, y = 1 - 2i/(n-1), radius sqrt(1 - y^2), angle
 * i * pi * (sqrt(5) - 1); Phi in double stored as float, the rest float
 * arithmetic (cosf / sinf / sqrtf), normalised as glm::normalize does:
 * v * (1 / sqrt((x*x + y*y) + z*z)). */
void mo_pool_bake_directions(int32_t n, float *dirs);
/* The baked direction nearest to the unit forward f (TNearestMap::Query,
 * Runtimes/Voxel/Spatial/NearestMap.h:36-50: Euclidean nearest; the first
 * on ties): the view the reference's BakeVisibilityByView
 * (ChunkManagerHelper.h:186-219) answers with. */
void mo_pool_snap(const float *dirs, int32_t n, const float f[3], float out[3]);
/* FImportanceComputeInfo::CalculateChunkImportance (ChunkManagerHelper.h:21-43) */
float mo_chunk_importance(int32_t dx, int32_t dy, int32_t dz, const float f[3]);
/* GetDesiredShowChunkLocationByView's acceptance (ChunkManagerHelper.h:98-135) */
int mo_chunk_desired(int32_t dx, int32_t dy, int32_t dz, const float f[3],
                     const mo_pool_params *p);
/* 1 when the chunk holds a non-empty brick (the scene's 32^3 level) */
int mo_chunk_nonempty(const mo_scene *s, int32_t cx, int32_t cy, int32_t cz);
/* One update.  Candidates: desired chunks inside the scene that are non-empty
 * and not resident, ordered by importance (descending), then chunk index.
 * Victims: the slots, free ones first (by slot), then resident chunks by
 * importance under this camera (ascending), then slot.  Candidate i takes
 * victim i while i < max_loads and the victim is free or strictly less
 * important (FindMin: a resident chunk is replaced only by a more important
 * one); the first failure ends the update (both orders are monotone).  Each
 * loaded chunk is meshed from the scene into its slot (quads at
 * slot * quads_per_slot, count in slot_count).  lookup: nc^3 entries (slot or
 * MO_NO_SLOT), slot_chunk / slot_stamp / slot_count: capacity entries;
 * load_xyz / load_slot: max_loads entries (the loads, in candidate order). */
void mo_pool_update(const mo_scene *s, const mo_pool_params *p, uint32_t stamp,
                    const int32_t cam[3], const float f[3], uint32_t *lookup,
                    uint32_t *slot_chunk, uint32_t *slot_stamp, uint32_t *slot_count,
                    uint32_t *quads, int32_t *load_xyz, uint32_t *load_slot,
                    mo_pool_result *res);

/* ---------------- quad visibility (the raster consumer of meshes) ---------------- */
/* Every pixel of cam gets the nearest front-facing quad of the pool's
 * resident slots along its primary ray (mo_primary_ray).  A quad of face f
 * lies on the plane P = chunk origin + slice (+1 for the positive faces
 * 0/2/4) of its axis; the ray hits it when it travels against the face
 * normal, t = (P - o[a]) * (1 / d[a]) > 0, and the hit point
 * h = o + d * t (per axis: d * t, then o + that) lies in [u0, u0 + w) x
 * [v0, v0 + h).  Per pixel the smallest key (float bits of t << 32) |
 * (slot * per_slot + k) wins.  Outputs are the raymarcher's: the voxel
 * behind the face (index x + n(y + n z); -1 on a miss), the face, t (+inf on
 * a miss) and the primary-only shade 0.25 + 0.75 * lambert (0 on a miss).
 * Which pixels a quad is tested on is a conservative screen rectangle (the
 * projected corners after clipping at 1e-3 in front of the eye, plus 1
 * pixel); the result depends only on the per-pixel test. */
void mo_raster_pool(const mo_camera *cam, int32_t n, int32_t nc, const uint32_t *slot_chunk,
                    const uint32_t *slot_count, const uint32_t *quads, int32_t capacity,
                    int32_t per_slot, const float light[3], uint64_t *keys, int64_t *out_index,
                    uint8_t *out_face, float *out_t, float *out_shade);

#ifdef __cplusplus
}
#endif
#endif
