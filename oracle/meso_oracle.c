/*
 * meso_oracle.c -- CPU oracle (test infrastructure only; see meso_oracle.h).
 *
 * Every float expression is written one IEEE operation at a time and the file
 * is built with -ffp-contract=off, so the results are the exact binary32
 * results of the operations as written.  The HIP library performs the same
 * operations in the same order; the tests compare the two bit for bit.
 */
#include "meso_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* conventions                                                          */
/* ------------------------------------------------------------------ */

uint32_t mo_quad_pack(int x, int y, int z, int w, int h, int face) {
  return (uint32_t)x | ((uint32_t)y << 5) | ((uint32_t)z << 10) |
         ((uint32_t)(w - 1) << 15) | ((uint32_t)(h - 1) << 20) | ((uint32_t)face << 25);
}

/* face id from the axis (0=x,1=y,2=z) and the sign of the outward normal:
 * Cube.h:25-54 lists +Z,-Z,+Y,-Y,+X,-X. */
static int face_id(int axis, int positive) {
  int base = axis == 2 ? 0 : (axis == 1 ? 2 : 4);
  return base + (positive ? 0 : 1);
}
/* face through which a ray stepping along `axis` with sign `step` enters a
 * voxel: stepping +x enters through the -X face. */
static int entry_face(int axis, int step) { return face_id(axis, step < 0); }
static int face_axis(int face) { return face < 2 ? 2 : (face < 4 ? 1 : 0); }
static int face_sign(int face) { return (face & 1) ? -1 : 1; }

/* Convert3DTo1D (VoxelMathHelper.h:73-76) for a cube of edge n */
static int64_t lin3(int64_t x, int64_t y, int64_t z, int64_t n) { return x + y * n + z * n * n; }

/* ------------------------------------------------------------------ */
/* RNG / hashing                                                        */
/* ------------------------------------------------------------------ */

/* SplitMix64 (S. Vigna, public domain reference algorithm) */
uint64_t mo_splitmix64(uint64_t *state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* PCG-XSH-RR 64/32 (M. O'Neill, pcg-random.org "minimal C" generator) */
uint32_t mo_pcg32_next(mo_pcg32 *r) {
  uint64_t old = r->state;
  r->state = old * 6364136223846793005ull + r->inc;
  uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
  uint32_t rot = (uint32_t)(old >> 59u);
  return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

void mo_pcg32_srandom(mo_pcg32 *r, uint64_t initstate, uint64_t initseq) {
  r->state = 0u;
  r->inc = (initseq << 1u) | 1u;
  mo_pcg32_next(r);
  r->state += initstate;
  mo_pcg32_next(r);
}

void mo_pcg32_seed(mo_pcg32 *r, uint64_t seed) {
  uint64_t sm = seed;
  uint64_t a = mo_splitmix64(&sm);
  uint64_t b = mo_splitmix64(&sm);
  mo_pcg32_srandom(r, a, b);
}

static float u01(mo_pcg32 *r) { return (float)(mo_pcg32_next(r) >> 8) * (1.0f / 16777216.0f); }

/* lowbias32 (C. Wellons, hash-prospector) */
uint32_t mo_hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

uint32_t mo_lattice_hash3(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  return mo_hash32((uint32_t)x ^ mo_hash32((uint32_t)y ^ mo_hash32((uint32_t)z ^ seed)));
}
static uint32_t lattice_hash2(int32_t x, int32_t z, uint32_t seed) {
  return mo_hash32((uint32_t)x ^ mo_hash32((uint32_t)z ^ seed));
}
static float hash_to_unit(uint32_t h) { return (float)(h >> 8) * (1.0f / 16777216.0f); }

/* quintic fade t^3 (t (6 t - 15) + 10) */
static float fade(float t) {
  float q = t * 6.0f;
  q = q - 15.0f;
  q = q * t;
  q = q + 10.0f;
  float t3 = t * t;
  t3 = t3 * t;
  return t3 * q;
}
static float lerpf(float a, float b, float t) {
  float d = b - a;
  d = d * t;
  return a + d;
}

float mo_value_noise2(float x, float z, uint32_t seed) {
  float fx = floorf(x), fz = floorf(z);
  int32_t ix = (int32_t)fx, iz = (int32_t)fz;
  float ux = fade(x - fx), uz = fade(z - fz);
  float v00 = hash_to_unit(lattice_hash2(ix, iz, seed));
  float v10 = hash_to_unit(lattice_hash2(ix + 1, iz, seed));
  float v01 = hash_to_unit(lattice_hash2(ix, iz + 1, seed));
  float v11 = hash_to_unit(lattice_hash2(ix + 1, iz + 1, seed));
  float a = lerpf(v00, v10, ux);
  float b = lerpf(v01, v11, ux);
  return lerpf(a, b, uz);
}

float mo_value_noise3(float x, float y, float z, uint32_t seed) {
  float fx = floorf(x), fy = floorf(y), fz = floorf(z);
  int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float ux = fade(x - fx), uy = fade(y - fy), uz = fade(z - fz);
  float v[8];
  for (int k = 0; k < 8; k++)
    v[k] = hash_to_unit(mo_lattice_hash3(ix + (k & 1), iy + ((k >> 1) & 1), iz + (k >> 2), seed));
  float x00 = lerpf(v[0], v[1], ux);
  float x10 = lerpf(v[2], v[3], ux);
  float x01 = lerpf(v[4], v[5], ux);
  float x11 = lerpf(v[6], v[7], ux);
  float y0 = lerpf(x00, x10, uy);
  float y1 = lerpf(x01, x11, uy);
  return lerpf(y0, y1, uz);
}

#define OCTAVE_SEED_STEP 0x9E3779B9u

float mo_fbm2(float x, float z, int octaves, float base_freq, uint32_t seed) {
  float sum = 0.0f, norm = 0.0f, amp = 1.0f, f = base_freq;
  for (int o = 0; o < octaves; o++) {
    float px = x * f, pz = z * f;
    float n = mo_value_noise2(px, pz, seed + (uint32_t)o * OCTAVE_SEED_STEP);
    float an = amp * n;
    sum = sum + an;
    norm = norm + amp;
    amp = amp * 0.5f;
    f = f * 2.0f;
  }
  return sum / norm;
}

float mo_fbm3(float x, float y, float z, int octaves, float base_freq, uint32_t seed) {
  float sum = 0.0f, norm = 0.0f, amp = 1.0f, f = base_freq;
  for (int o = 0; o < octaves; o++) {
    float px = x * f, py = y * f, pz = z * f;
    float n = mo_value_noise3(px, py, pz, seed + (uint32_t)o * OCTAVE_SEED_STEP);
    float an = amp * n;
    sum = sum + an;
    norm = norm + amp;
    amp = amp * 0.5f;
    f = f * 2.0f;
  }
  return sum / norm;
}

/* ------------------------------------------------------------------ */
/* generators                                                           */
/* ------------------------------------------------------------------ */

void mo_sphere_params(uint64_t seed, int n, int n_spheres, int32_t *params) {
  mo_pcg32 r;
  mo_pcg32_seed(&r, seed);
  uint32_t span = (uint32_t)n * 256u;
  for (int i = 0; i < n_spheres; i++) {
    params[4 * i + 0] = (int32_t)(mo_pcg32_next(&r) % span);
    params[4 * i + 1] = (int32_t)(mo_pcg32_next(&r) % span);
    params[4 * i + 2] = (int32_t)(mo_pcg32_next(&r) % span);
    params[4 * i + 3] = (int32_t)(8 * 256 + mo_pcg32_next(&r) % (24u * 256u + 1u));
  }
}

/* voxel centre (x + 1/2) in 1/256 fixed point against sphere i */
static int in_sphere_fp(const int32_t *s, int32_t x, int32_t y, int32_t z) {
  int64_t dx = (int64_t)x * 256 + 128 - s[0];
  int64_t dy = (int64_t)y * 256 + 128 - s[1];
  int64_t dz = (int64_t)z * 256 + 128 - s[2];
  int64_t r = s[3];
  return dx * dx + dy * dy + dz * dz <= r * r;
}

void mo_gen_spheres_dense(const int32_t *params, int n_spheres, int n, uint8_t *grid) {
  for (int z = 0; z < n; z++)
    for (int y = 0; y < n; y++)
      for (int x = 0; x < n; x++) {
        uint8_t m = 0;
        for (int i = 0; i < n_spheres; i++)
          if (in_sphere_fp(params + 4 * i, x, y, z)) { m = (uint8_t)(1 + i); break; }
        grid[lin3(x, y, z, n)] = m;
      }
}

int32_t mo_height(int32_t x, int32_t z, int32_t n, uint32_t seed) {
  float f = mo_fbm2((float)x, (float)z, 4, 4.0f / (float)n, seed);
  f = f - 0.5f;
  float h = 0.40f * f;
  h = 0.10f + h;
  h = (float)n * h;
  return (int32_t)floorf(h);
}

static uint16_t height_band(int32_t y, int32_t n) {
  int32_t b = (y * 16) / n;
  return (uint16_t)(1 + (b < 3 ? b : 3));
}

uint16_t mo_heightfield_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed) {
  if (y >= mo_height(x, z, n, seed)) return 0;
  return height_band(y, n);
}

static int32_t terrain_height(int32_t x, int32_t z, int32_t n, uint32_t seed) {
  float f = mo_fbm2((float)x, (float)z, 4, 4.0f / (float)n, seed);
  float h = 0.40f * f;
  h = 0.05f + h;
  h = (float)n * h;
  return (int32_t)floorf(h);
}

uint16_t mo_terrain_caves_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed) {
  if (y >= terrain_height(x, z, n, seed)) return 0;
  float c = mo_fbm3((float)x, (float)y, (float)z, 3, 8.0f / (float)n, seed ^ 0x5bd1e995u);
  c = c - 0.5f;
  if (fabsf(c) < 0.03f) return 0; /* tunnel */
  return height_band(y, n);
}

uint16_t mo_caves_voxel(int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed) {
  (void)n;
  float c = mo_fbm3((float)x, (float)y, (float)z, 4, 1.0f / 64.0f, seed);
  c = c - 0.5f;
  if (fabsf(c) >= 0.0055f) return 0;
  return (uint16_t)(1 + ((uint32_t)y >> 7) % 4u);
}

uint16_t mo_scene_voxel(int kind, int32_t x, int32_t y, int32_t z, int32_t n, uint32_t seed) {
  switch (kind) {
    case MO_GEN_HEIGHTFIELD: return mo_heightfield_voxel(x, y, z, n, seed);
    case MO_GEN_TERRAIN_CAVES: return mo_terrain_caves_voxel(x, y, z, n, seed);
    case MO_GEN_CAVES: return mo_caves_voxel(x, y, z, n, seed);
    default: return 0;
  }
}

int32_t mo_gen_brick(int kind, int32_t n, uint32_t seed, int32_t bx, int32_t by, int32_t bz,
                     uint16_t *vox, uint64_t *mask) {
  int32_t solid = 0;
  for (int k = 0; k < 8; k++) mask[k] = 0;
  for (int lz = 0; lz < 8; lz++)
    for (int ly = 0; ly < 8; ly++)
      for (int lx = 0; lx < 8; lx++) {
        uint16_t v = mo_scene_voxel(kind, bx * 8 + lx, by * 8 + ly, bz * 8 + lz, n, seed);
        vox[lx + 8 * ly + 64 * lz] = v;
        if (v) {
          mask[lz] |= 1ull << (lx + 8 * ly);
          solid++;
        }
      }
  return solid;
}

int64_t mo_build_brickmap(int kind, int32_t n, uint32_t seed, uint32_t *slots, uint64_t *masks,
                          uint16_t *voxels, int64_t cap) {
  int32_t nb = n / 8;
  int32_t *heights = NULL;
  if (kind == MO_GEN_HEIGHTFIELD) {
    heights = (int32_t *)malloc(sizeof(int32_t) * (size_t)n * (size_t)n);
    for (int32_t z = 0; z < n; z++)
      for (int32_t x = 0; x < n; x++) heights[x + (int64_t)n * z] = mo_height(x, z, n, seed);
  }
  uint16_t vox[512];
  uint64_t mk[8];
  int64_t count = 0;
  for (int32_t bz = 0; bz < nb; bz++)
    for (int32_t by = 0; by < nb; by++)
      for (int32_t bx = 0; bx < nb; bx++) {
        int32_t solid = 0;
        if (heights) {
          for (int k = 0; k < 8; k++) mk[k] = 0;
          for (int lz = 0; lz < 8; lz++)
            for (int ly = 0; ly < 8; ly++)
              for (int lx = 0; lx < 8; lx++) {
                int32_t x = bx * 8 + lx, y = by * 8 + ly, z = bz * 8 + lz;
                uint16_t v = y < heights[x + (int64_t)n * z] ? height_band(y, n) : 0;
                vox[lx + 8 * ly + 64 * lz] = v;
                if (v) { mk[lz] |= 1ull << (lx + 8 * ly); solid++; }
              }
        } else {
          solid = mo_gen_brick(kind, n, seed, bx, by, bz, vox, mk);
        }
        int64_t b = lin3(bx, by, bz, nb);
        if (solid == 0) {
          if (slots) slots[b] = MO_EMPTY_SLOT;
          continue;
        }
        if (slots) slots[b] = (uint32_t)count;
        if (masks && count < cap) {
          memcpy(masks + 8 * count, mk, 64);
          memcpy(voxels + 512 * count, vox, 1024);
        }
        count++;
      }
  free(heights);
  return count;
}

int64_t mo_brickmap_from_dense(const uint16_t *grid, int32_t n, uint32_t *slots, uint64_t *masks,
                               uint16_t *voxels, int64_t cap) {
  int32_t nb = n / 8;
  int64_t count = 0;
  uint16_t vox[512];
  uint64_t mk[8];
  for (int32_t bz = 0; bz < nb; bz++)
    for (int32_t by = 0; by < nb; by++)
      for (int32_t bx = 0; bx < nb; bx++) {
        int32_t solid = 0;
        for (int k = 0; k < 8; k++) mk[k] = 0;
        for (int lz = 0; lz < 8; lz++)
          for (int ly = 0; ly < 8; ly++)
            for (int lx = 0; lx < 8; lx++) {
              uint16_t v = grid[lin3(bx * 8 + lx, by * 8 + ly, bz * 8 + lz, n)];
              vox[lx + 8 * ly + 64 * lz] = v;
              if (v) { mk[lz] |= 1ull << (lx + 8 * ly); solid++; }
            }
        int64_t b = lin3(bx, by, bz, nb);
        if (!solid) { slots[b] = MO_EMPTY_SLOT; continue; }
        slots[b] = (uint32_t)count;
        if (masks && count < cap) {
          memcpy(masks + 8 * count, mk, 64);
          memcpy(voxels + 512 * count, vox, 1024);
        }
        count++;
      }
  return count;
}

/* chunk batch: Beta(1/2,1/2) fill as cos^2 of a uniform angle, the angle
 * taken from a point drawn uniformly in the unit disk by rejection. */
void mo_chunk_params_gen(uint64_t seed, int32_t n_chunks, mo_chunk_params *out) {
  mo_pcg32 r;
  mo_pcg32_seed(&r, seed);
  for (int32_t c = 0; c < n_chunks; c++) {
    mo_chunk_params *p = out + c;
    float a, b, r2;
    for (;;) {
      a = u01(&r) * 2.0f;
      a = a - 1.0f;
      b = u01(&r) * 2.0f;
      b = b - 1.0f;
      float aa = a * a, bb = b * b;
      r2 = aa + bb;
      if (r2 < 1.0f && r2 > 0.0f) break;
    }
    float aa = a * a;
    p->fill = aa / r2;
    p->noise_seed = mo_pcg32_next(&r);
    p->n_spheres = 1 + (int32_t)(mo_pcg32_next(&r) % 4u);
    for (int i = 0; i < 4; i++) {
      p->sph[i][0] = (int32_t)(mo_pcg32_next(&r) % (32u * 256u));
      p->sph[i][1] = (int32_t)(mo_pcg32_next(&r) % (32u * 256u));
      p->sph[i][2] = (int32_t)(mo_pcg32_next(&r) % (32u * 256u));
      p->sph[i][3] = (int32_t)(3 * 256 + mo_pcg32_next(&r) % (9u * 256u + 1u));
    }
  }
}

static int32_t chunk_noise_q(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  float v = mo_fbm3((float)x, (float)y, (float)z, 2, 0.125f, seed);
  v = v * 256.0f;
  int32_t q = (int32_t)v;
  return q > 255 ? 255 : q;
}

void mo_gen_chunk(const mo_chunk_params *p, uint16_t *vox) {
  uint32_t hist[256];
  memset(hist, 0, sizeof(hist));
  uint8_t *q = (uint8_t *)malloc(32768);
  for (int32_t z = 0; z < 32; z++)
    for (int32_t y = 0; y < 32; y++)
      for (int32_t x = 0; x < 32; x++) {
        int32_t v = chunk_noise_q(x, y, z, p->noise_seed);
        q[x + 32 * y + 1024 * z] = (uint8_t)v;
        hist[v]++;
      }
  float tf = p->fill * 32768.0f;
  tf = tf + 0.5f;
  uint32_t target = (uint32_t)tf;
  uint32_t thr = 0, cum = 0;
  while (cum < target && thr < 256) { cum += hist[thr]; thr++; }
  for (int32_t z = 0; z < 32; z++)
    for (int32_t y = 0; y < 32; y++)
      for (int32_t x = 0; x < 32; x++) {
        int32_t i = x + 32 * y + 1024 * z;
        uint16_t m = 0;
        for (int s = 0; s < p->n_spheres; s++)
          if (in_sphere_fp(p->sph[s], x, y, z)) { m = (uint16_t)(5 + s); break; }
        if (!m && (uint32_t)q[i] < thr) m = (uint16_t)(1 + ((uint32_t)q[i] * 4u) / thr);
        vox[i] = m;
      }
  free(q);
}

/* ------------------------------------------------------------------ */
/* camera                                                               */
/* ------------------------------------------------------------------ */

static void normalize3(float v[3]) {
  float xx = v[0] * v[0], yy = v[1] * v[1], zz = v[2] * v[2];
  float s = xx + yy;
  s = s + zz;
  float len = sqrtf(s);
  v[0] = v[0] / len;
  v[1] = v[1] / len;
  v[2] = v[2] / len;
}
static void cross3(const float a[3], const float b[3], float out[3]) {
  float t0 = a[1] * b[2], t1 = a[2] * b[1];
  float t2 = a[2] * b[0], t3 = a[0] * b[2];
  float t4 = a[0] * b[1], t5 = a[1] * b[0];
  out[0] = t0 - t1;
  out[1] = t2 - t3;
  out[2] = t4 - t5;
}

/* vertical-FOV pinhole as FVoxelCamera::GetProjectionMatrix builds with
 * glm::perspective (Runtimes/Instance/VoxelCamera.cpp:12-17); forward points
 * from the eye to the target (VoxelCamera.cpp:24-29 takes -row 2 of view). */
void mo_camera_derive(const float eye[3], const float target[3], const float up[3], float vfov_deg,
                      int32_t width, int32_t height, mo_camera *c) {
  for (int i = 0; i < 3; i++) {
    c->eye[i] = eye[i];
    c->fwd[i] = target[i] - eye[i];
  }
  normalize3(c->fwd);
  cross3(c->fwd, up, c->right);
  normalize3(c->right);
  cross3(c->right, c->fwd, c->up);
  float th = (float)tan((double)vfov_deg * 3.14159265358979323846 / 360.0);
  float aspect = (float)width / (float)height;
  c->sx = th * aspect;
  c->sy = th;
  c->inv_w2 = 2.0f / (float)width;
  c->inv_h2 = 2.0f / (float)height;
  c->width = width;
  c->height = height;
}

void mo_primary_ray(const mo_camera *c, int32_t px, int32_t py, float o[3], float d[3]) {
  float a = (float)px + 0.5f;
  a = a * c->inv_w2;
  a = a - 1.0f;
  a = a * c->sx;
  float b = (float)py + 0.5f;
  b = b * c->inv_h2;
  b = 1.0f - b;
  b = b * c->sy;
  float v[3];
  for (int i = 0; i < 3; i++) {
    float ra = c->right[i] * a;
    float ub = c->up[i] * b;
    v[i] = c->fwd[i] + ra;
    v[i] = v[i] + ub;
  }
  for (int i = 0; i < 3; i++) o[i] = c->eye[i];
  float xx = v[0] * v[0], yy = v[1] * v[1], zz = v[2] * v[2];
  float s = xx + yy;
  s = s + zz;
  float len = sqrtf(s);
  for (int i = 0; i < 3; i++) d[i] = v[i] / len;
}

/* ------------------------------------------------------------------ */
/* traversal                                                            */
/* ------------------------------------------------------------------ */

typedef struct {
  float o[3], d[3], inv[3];
  int32_t step[3];
} ray_t;

static void ray_setup(ray_t *r, const float o[3], const float d[3]) {
  for (int a = 0; a < 3; a++) {
    r->o[a] = o[a];
    r->d[a] = d[a];
    if (d[a] > 0.0f) { r->step[a] = 1; r->inv[a] = 1.0f / d[a]; }
    else if (d[a] < 0.0f) { r->step[a] = -1; r->inv[a] = 1.0f / d[a]; }
    else { r->step[a] = 0; r->inv[a] = 0.0f; }
  }
}
static float plane_t(const ray_t *r, int a, float b) {
  float q = b - r->o[a];
  return q * r->inv[a];
}
/* smallest t; ties go to x, then y (the x<y<z rule) */
static int argmin3(const float t[3]) {
  if (t[0] <= t[1] && t[0] <= t[2]) return 0;
  return t[1] <= t[2] ? 1 : 2;
}

typedef struct {
  int32_t cell[3];
  float t;
  int32_t face;
} walk_state;

static int get_bit(const uint32_t *bits, int64_t i) { return (int)((bits[i >> 5] >> (i & 31)) & 1u); }
static int32_t ceil_div(int32_t a, int32_t b) { return (a + b - 1) / b; }

/* One traversal step at level L (1 = voxel, 8 = brick, 32, 128 = empty
 * blocks): leave the block of edge L that holds the current cell through its
 * nearest plane (ties x, then y).  For L > 1 the other two coordinates come
 * from the exit point, clamped into the block (and into the scene) and never
 * moving against the ray; for L = 1 they are unchanged, so the step is the
 * plain Amanatides & Woo voxel step.  Returns 0 when the step along the
 * chosen axis leaves the scene's box (nothing solid lies beyond it). */
static int advance(const ray_t *r, int32_t n, const int32_t *box_lo, const int32_t *box_hi,
                   int32_t L, walk_state *w) {
  int32_t bc[3];
  float tb[3];
  for (int a = 0; a < 3; a++) {
    bc[a] = w->cell[a] / L;
    if (r->step[a] > 0) tb[a] = plane_t(r, a, (float)((bc[a] + 1) * L));
    else if (r->step[a] < 0) tb[a] = plane_t(r, a, (float)(bc[a] * L));
    else tb[a] = INFINITY;
  }
  int a = argmin3(tb);
  float t = tb[a];
  int32_t na = r->step[a] > 0 ? (bc[a] + 1) * L : bc[a] * L - 1;
  if (na < box_lo[a] || na >= box_hi[a]) return 0;
  if (L > 1) {
    for (int b = 0; b < 3; b++) {
      if (b == a) continue;
      float dt = r->d[b] * t;
      float p = r->o[b] + dt;
      int32_t c = (int32_t)floorf(p);
      int32_t lo = bc[b] * L, hi = lo + L - 1;
      if (hi > n - 1) hi = n - 1;
      if (c < lo) c = lo;
      if (c > hi) c = hi;
      if (r->step[b] > 0 && c < w->cell[b]) c = w->cell[b];
      if (r->step[b] < 0 && c > w->cell[b]) c = w->cell[b];
      w->cell[b] = c;
    }
  }
  w->cell[a] = na;
  w->t = t;
  w->face = entry_face(a, r->step[a]);
  return 1;
}

static int same3(const int32_t a[3], const int32_t b[3]) {
  return a[0] == b[0] && a[1] == b[1] && a[2] == b[2];
}

/* 1 if the aligned B^3 block (B = 2 or 4) of the brick that holds voxel
 * (lx, ly, lz) has no solid voxel (mask word z, bit x + 8 y). */
static int brick_block_empty(const uint64_t *mask, int32_t lx, int32_t ly, int32_t lz, int32_t B) {
  const int32_t x0 = lx & ~(B - 1), y0 = ly & ~(B - 1), z0 = lz & ~(B - 1);
  for (int32_t z = z0; z < z0 + B; z++)
    for (int32_t y = y0; y < y0 + B; y++)
      for (int32_t x = x0; x < x0 + B; x++)
        if ((mask[z] >> (x + 8 * y)) & 1u) return 0;
  return 1;
}

/* Walk from state w until a solid voxel is entered (returns 1), the ray
 * leaves the scene or the entry t reaches tmax (returns 0).  Each step picks
 * the coarsest empty level holding the current cell: 128^3 block, 32^3
 * block, brick, 4^3 and 2^3 blocks inside a non-empty brick (read from its
 * mask), else the voxel itself. */
static int walk(const mo_scene *s, const ray_t *r, walk_state *w, float tmax, uint16_t *material,
                mo_counters *cnt) {
  const int32_t n = s->n;
  const int two = s->kind == MO_SCENE_TWO_LEVEL_U16;
  const int32_t n128 = ceil_div(n, 128), n32 = ceil_div(n, 32);
  int32_t c128[3] = {-1, -1, -1}, c32[3] = {-1, -1, -1}, cb[3] = {-1, -1, -1};
  int e128 = 0, e32 = 0;
  uint32_t node = MO_EMPTY_SLOT, slot = MO_EMPTY_SLOT;
  uint64_t groups = 0;
  for (;;) {
    if (!(w->t < tmax)) return 0;
    cnt->n_iter++;
    int32_t L = 1;
    if (s->kind == MO_SCENE_DENSE_U8) {
      cnt->n_vox++;
      uint8_t v = s->dense[lin3(w->cell[0], w->cell[1], w->cell[2], n)];
      if (v) { *material = v; return 1; }
    } else {
      int32_t k[3] = {w->cell[0] >> 7, w->cell[1] >> 7, w->cell[2] >> 7};
      if (!same3(k, c128)) {
        memcpy(c128, k, sizeof(k));
        c32[0] = c32[1] = c32[2] = -1;
        cb[0] = cb[1] = cb[2] = -1;
        if (two) {
          cnt->n_top++;
          node = s->top[lin3(k[0], k[1], k[2], n128)];
          e128 = node == MO_EMPTY_SLOT;
          groups = e128 ? 0 : s->groups[node];
        } else {
          cnt->n_coarse++;
          e128 = !get_bit(s->occ128, lin3(k[0], k[1], k[2], n128));
        }
      }
      if (e128) {
        L = 128;
      } else {
        int32_t g[3] = {w->cell[0] >> 5, w->cell[1] >> 5, w->cell[2] >> 5};
        if (!same3(g, c32)) {
          memcpy(c32, g, sizeof(g));
          cb[0] = cb[1] = cb[2] = -1;
          if (two) {
            cnt->n_coarse++; /* the 32^3 block's word of allocated bricks */
            e32 = !((groups >> ((g[0] & 3) + 4 * (g[1] & 3) + 16 * (g[2] & 3))) & 1u);
          } else {
            cnt->n_coarse++;
            e32 = !get_bit(s->occ32, lin3(g[0], g[1], g[2], n32));
          }
        }
        if (e32) {
          L = 32;
        } else {
          int32_t bc[3] = {w->cell[0] >> 3, w->cell[1] >> 3, w->cell[2] >> 3};
          if (!same3(bc, cb)) {
            memcpy(cb, bc, sizeof(bc));
            if (two) {
              cnt->n_node++;
              slot = s->slots[(int64_t)node * 4096 + lin3(bc[0] & 15, bc[1] & 15, bc[2] & 15, 16)];
            } else {
              cnt->n_brickslot++;
              slot = s->slots[lin3(bc[0], bc[1], bc[2], n >> 3)];
            }
            if (slot != MO_EMPTY_SLOT) cnt->n_brickmask++;
          }
          if (slot == MO_EMPTY_SLOT) {
            L = 8;
          } else {
            int32_t lx = w->cell[0] & 7, ly = w->cell[1] & 7, lz = w->cell[2] & 7;
            const uint64_t *mask = s->masks + (int64_t)slot * 8;
            if (brick_block_empty(mask, lx, ly, lz, 4)) {
              L = 4;
            } else if (brick_block_empty(mask, lx, ly, lz, 2)) {
              L = 2;
            } else {
              cnt->n_vox++; /* one 8-byte mask word read */
              if ((mask[lz] >> (lx + 8 * ly)) & 1u) {
                *material = s->voxels[(int64_t)slot * 512 + lx + 8 * ly + 64 * lz];
                return 1;
              }
            }
          }
        }
      }
    }
    if (!advance(r, n, s->box_lo, s->box_hi, L, w)) return 0;
  }
}

void mo_build_coarse(const mo_scene *s, int64_t n_nodes, uint32_t *occ32, uint32_t *occ128,
                     uint64_t *groups) {
  const int32_t n = s->n, nb = n / 8;
  if (s->kind == MO_SCENE_BRICKMAP_U16) {
    const int32_t n32 = ceil_div(n, 32), n128 = ceil_div(n, 128);
    memset(occ128, 0, sizeof(uint32_t) * (size_t)(((int64_t)n128 * n128 * n128 + 31) / 32));
    memset(occ32, 0, sizeof(uint32_t) * (size_t)(((int64_t)n32 * n32 * n32 + 31) / 32));
    for (int32_t bz = 0; bz < nb; bz++)
      for (int32_t by = 0; by < nb; by++)
        for (int32_t bx = 0; bx < nb; bx++) {
          if (s->slots[lin3(bx, by, bz, nb)] == MO_EMPTY_SLOT) continue;
          int64_t i32 = lin3(bx >> 2, by >> 2, bz >> 2, n32);
          int64_t i128 = lin3(bx >> 4, by >> 4, bz >> 4, n128);
          occ32[i32 >> 5] |= 1u << (i32 & 31);
          occ128[i128 >> 5] |= 1u << (i128 & 31);
        }
  } else if (s->kind == MO_SCENE_TWO_LEVEL_U16) {
    for (int64_t node = 0; node < n_nodes; node++) {
      uint64_t g = 0;
      for (int32_t l = 0; l < 4096; l++) {
        if (s->slots[node * 4096 + l] == MO_EMPTY_SLOT) continue;
        int32_t bx = l & 15, by = (l >> 4) & 15, bz = l >> 8;
        g |= 1ull << ((bx >> 2) + 4 * (by >> 2) + 16 * (bz >> 2));
      }
      groups[node] = g;
    }
  }
}

int mo_trace(const mo_scene *s, const float o[3], const float d[3], mo_hit *hit, mo_counters *cnt) {
  ray_t r;
  ray_setup(&r, o, d);
  const int32_t n = s->n;
  hit->index = -1;
  hit->face = MO_FACE_MISS;
  hit->t = INFINITY;
  hit->material = 0;
  /* clip to the scene's box */
  float t_enter = 0.0f, t_exit = INFINITY;
  int ea = -1;
  for (int a = 0; a < 3; a++) {
    if (s->box_lo[a] >= s->box_hi[a]) return 0; /* nothing solid */
    const float bl = (float)s->box_lo[a], bh = (float)s->box_hi[a];
    if (r.step[a] == 0) {
      if (o[a] < bl || o[a] >= bh) return 0;
      continue;
    }
    float lo = plane_t(&r, a, r.step[a] > 0 ? bl : bh);
    float hi = plane_t(&r, a, r.step[a] > 0 ? bh : bl);
    if (lo > t_enter) { t_enter = lo; ea = a; }
    if (hi < t_exit) t_exit = hi;
  }
  if (t_enter > t_exit) return 0;
  walk_state w;
  for (int a = 0; a < 3; a++) {
    if (a == ea) {
      w.cell[a] = r.step[a] > 0 ? s->box_lo[a] : s->box_hi[a] - 1;
    } else {
      float dt = d[a] * t_enter;
      float p = o[a] + dt;
      int32_t c = (int32_t)floorf(p);
      if (c < s->box_lo[a]) c = s->box_lo[a];
      if (c > s->box_hi[a] - 1) c = s->box_hi[a] - 1;
      w.cell[a] = c;
    }
  }
  w.t = t_enter;
  w.face = ea >= 0 ? entry_face(ea, r.step[ea]) : MO_FACE_INSIDE;
  uint16_t m = 0;
  if (!walk(s, &r, &w, INFINITY, &m, cnt)) return 0;
  cnt->hits++;
  hit->index = lin3(w.cell[0], w.cell[1], w.cell[2], n);
  hit->face = w.face;
  hit->t = w.t;
  hit->material = m;
  memcpy(hit->cell, w.cell, sizeof(w.cell));
  return 1;
}

int mo_occluded(const mo_scene *s, const float o[3], const float d[3], const int32_t start[3],
                float tmax, mo_counters *cnt) {
  /* a start outside the box (the neighbour across a face of a voxel on the
   * box's side) sees nothing: the ray travels into the face's half-space,
   * away from the box */
  for (int a = 0; a < 3; a++)
    if (start[a] < s->box_lo[a] || start[a] >= s->box_hi[a]) return 0;
  ray_t r;
  ray_setup(&r, o, d);
  walk_state w;
  memcpy(w.cell, start, sizeof(w.cell));
  w.t = 0.0f;
  w.face = MO_FACE_INSIDE;
  uint16_t m;
  return walk(s, &r, &w, tmax, &m, cnt);
}

void mo_ao_direction(uint32_t seed, uint32_t pixel, uint32_t sample, int face, float dir[3]) {
  float a = 0.0f, b = 0.0f, r2 = 0.0f;
  uint32_t base = mo_hash32(mo_hash32(seed) ^ pixel);
  int ok = 0;
  for (uint32_t k = 0; k < 16u; k++) {
    uint32_t h = mo_hash32(base ^ (sample * 16u + k));
    a = (float)(h & 0xFFFFu) * (1.0f / 32768.0f);
    a = a - 1.0f;
    b = (float)(h >> 16) * (1.0f / 32768.0f);
    b = b - 1.0f;
    float aa = a * a, bb = b * b;
    r2 = aa + bb;
    if (r2 < 1.0f) { ok = 1; break; }
  }
  if (!ok) { a = 0.0f; b = 0.0f; r2 = 0.0f; }
  float z = sqrtf(1.0f - r2);
  int na = face_axis(face);
  int ua = na == 0 ? 1 : 0;
  int va = na == 2 ? 1 : 2;
  dir[na] = face_sign(face) > 0 ? z : -z;
  dir[ua] = a;
  dir[va] = b;
}

void mo_raymarch(const mo_scene *s, const mo_camera *cam, uint32_t flags, const mo_shading *sh,
                 int32_t row0, int32_t row1, int64_t *out_index, uint8_t *out_face, float *out_t,
                 float *out_shade, mo_counters *cnt) {
  const int32_t W = cam->width;
  for (int32_t py = row0; py < row1; py++)
    for (int32_t px = 0; px < W; px++) {
      int64_t pix = (int64_t)py * W + px;
      float o[3], d[3];
      mo_primary_ray(cam, px, py, o, d);
      mo_hit h;
      cnt->rays++;
      int hit = mo_trace(s, o, d, &h, cnt);
      float shade = 0.0f;
      if (hit) {
        if (h.face < 6) {
          int na = face_axis(h.face), sg = face_sign(h.face);
          float lambert = sg > 0 ? sh->light[na] : -sh->light[na];
          if (lambert < 0.0f) lambert = 0.0f;
          float p[3];
          for (int i = 0; i < 3; i++) {
            float dt = d[i] * h.t;
            p[i] = o[i] + dt;
          }
          int32_t start[3] = {h.cell[0], h.cell[1], h.cell[2]};
          start[na] += sg;
          float vis = 1.0f;
          if ((flags & MO_RAY_SHADOW) && lambert > 0.0f) {
            cnt->shadow_rays++;
            if (mo_occluded(s, p, sh->light, start, INFINITY, cnt)) vis = 0.0f;
          }
          float ao = 1.0f;
          if (flags & MO_RAY_AO) {
            int32_t unocc = 0;
            for (int32_t k = 0; k < sh->ao_samples; k++) {
              float dir[3];
              mo_ao_direction(sh->ao_seed, (uint32_t)pix, (uint32_t)k, h.face, dir);
              cnt->ao_rays++;
              if (!mo_occluded(s, p, dir, start, sh->ao_radius, cnt)) unocc++;
            }
            float frac = (float)unocc / (float)sh->ao_samples;
            frac = 0.5f * frac;
            ao = 0.5f + frac;
          }
          float l = lambert * vis;
          l = 0.75f * l;
          shade = 0.25f + l;
          shade = shade * ao;
        } else {
          shade = 0.25f;
        }
      }
      if (out_index) out_index[pix] = h.index;
      if (out_face) out_face[pix] = (uint8_t)h.face;
      if (out_t) out_t[pix] = h.t;
      if (out_shade) out_shade[pix] = shade;
    }
}

typedef struct {
  const mo_scene *s;
  const mo_camera *cam;
  uint32_t flags;
  const mo_shading *sh;
  int32_t row0, row1;
  int64_t *idx;
  uint8_t *face;
  float *t, *shade;
  mo_counters cnt;
} rm_job;

static void *rm_worker(void *p) {
  rm_job *j = (rm_job *)p;
  memset(&j->cnt, 0, sizeof(j->cnt));
  mo_raymarch(j->s, j->cam, j->flags, j->sh, j->row0, j->row1, j->idx, j->face, j->t, j->shade,
              &j->cnt);
  return NULL;
}

void mo_raymarch_mt(const mo_scene *s, const mo_camera *cam, uint32_t flags, const mo_shading *sh,
                    int32_t row0, int32_t row1, int32_t nthreads, int64_t *out_index,
                    uint8_t *out_face, float *out_t, float *out_shade, mo_counters *cnt) {
  if (nthreads < 1) nthreads = 1;
  rm_job *jobs = (rm_job *)calloc((size_t)nthreads, sizeof(rm_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  int32_t rows = row1 - row0;
  for (int32_t i = 0; i < nthreads; i++) {
    jobs[i] = (rm_job){s, cam, flags, sh, row0 + (int32_t)((int64_t)rows * i / nthreads),
                       row0 + (int32_t)((int64_t)rows * (i + 1) / nthreads),
                       out_index, out_face, out_t, out_shade, {0}};
    pthread_create(&th[i], NULL, rm_worker, &jobs[i]);
  }
  for (int32_t i = 0; i < nthreads; i++) {
    pthread_join(th[i], NULL);
    cnt->n_vox += jobs[i].cnt.n_vox;
    cnt->n_top += jobs[i].cnt.n_top;
    cnt->n_node += jobs[i].cnt.n_node;
    cnt->n_brickslot += jobs[i].cnt.n_brickslot;
    cnt->n_brickmask += jobs[i].cnt.n_brickmask;
    cnt->hits += jobs[i].cnt.hits;
    cnt->rays += jobs[i].cnt.rays;
    cnt->shadow_rays += jobs[i].cnt.shadow_rays;
    cnt->ao_rays += jobs[i].cnt.ao_rays;
    cnt->n_coarse += jobs[i].cnt.n_coarse;
    cnt->n_iter += jobs[i].cnt.n_iter;
  }
  free(jobs);
  free(th);
}

/* ------------------------------------------------------------------ */
/* meshing                                                              */
/* ------------------------------------------------------------------ */

static uint16_t chunk_get(const void *vox, int es, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= 32 || y >= 32 || z >= 32) return 0;
  int i = x + 32 * y + 1024 * z;
  return es == 1 ? ((const uint8_t *)vox)[i] : ((const uint16_t *)vox)[i];
}

/* plane coordinates: z faces (u,v)=(x,y); y faces (x,z); x faces (y,z) */
static void plane_to_xyz(int na, int s, int u, int v, int *x, int *y, int *z) {
  if (na == 2) { *x = u; *y = v; *z = s; }
  else if (na == 1) { *x = u; *y = s; *z = v; }
  else { *x = s; *y = u; *z = v; }
}

/* the plane's quads in output order: by last row (v + h - 1), then start u */
typedef struct {
  uint16_t key; /* (v + h - 1) * 32 + u */
  uint32_t quad;
} plane_quad;

static int plane_quad_cmp(const void *a, const void *b) {
  return (int)((const plane_quad *)a)->key - (int)((const plane_quad *)b)->key;
}

uint32_t mo_mesh_chunk(const void *vox, int es, int greedy, uint32_t *quads, uint32_t cap) {
  uint32_t count = 0;
  uint16_t mat[32][32];
  uint8_t open[32][32];
  plane_quad pq[1024];
  for (int f = 0; f < 6; f++) {
    int na = face_axis(f), sg = face_sign(f);
    for (int s = 0; s < 32; s++) {
      for (int v = 0; v < 32; v++)
        for (int u = 0; u < 32; u++) {
          int x, y, z;
          plane_to_xyz(na, s, u, v, &x, &y, &z);
          uint16_t m = chunk_get(vox, es, x, y, z);
          int nx = x, ny = y, nz = z;
          if (na == 0) nx += sg; else if (na == 1) ny += sg; else nz += sg;
          open[v][u] = m != 0 && chunk_get(vox, es, nx, ny, nz) == 0;
          mat[v][u] = m;
        }
      int np = 0;
      for (int v = 0; v < 32; v++)
        for (int u = 0; u < 32; u++) {
          if (!open[v][u]) continue;
          int w = 1, h = 1;
          uint16_t m = mat[v][u];
          if (greedy) {
            while (u + w < 32 && open[v][u + w] && mat[v][u + w] == m) w++;
            for (;;) {
              if (v + h >= 32) break;
              int ok = 1;
              for (int k = 0; k < w; k++)
                if (!open[v + h][u + k] || mat[v + h][u + k] != m) { ok = 0; break; }
              if (!ok) break;
              h++;
            }
          }
          for (int j = 0; j < h; j++)
            for (int k = 0; k < w; k++) open[v + j][u + k] = 0;
          int x, y, z;
          plane_to_xyz(na, s, u, v, &x, &y, &z);
          pq[np].key = (uint16_t)((v + h - 1) * 32 + u);
          pq[np].quad = mo_quad_pack(x, y, z, w, h, f);
          np++;
        }
      qsort(pq, (size_t)np, sizeof(plane_quad), plane_quad_cmp);  /* keys are distinct */
      for (int i = 0; i < np; i++) {
        if (count < cap && quads) quads[count] = pq[i].quad;
        count++;
      }
    }
  }
  return count;
}

typedef struct {
  const uint8_t *vox;
  int es, greedy;
  int32_t c0, c1;
  uint32_t *counts;
  uint64_t total;
} mesh_job;

static void *mesh_worker(void *p) {
  mesh_job *j = (mesh_job *)p;
  j->total = 0;
  for (int32_t c = j->c0; c < j->c1; c++) {
    uint32_t k = mo_mesh_chunk(j->vox + (size_t)c * 32768u * (size_t)j->es, j->es, j->greedy, NULL, 0);
    if (j->counts) j->counts[c] = k;
    j->total += k;
  }
  return NULL;
}

uint64_t mo_mesh_chunks_mt(const void *vox, int es, int32_t n_chunks, int greedy, int32_t nthreads,
                           uint32_t *counts) {
  if (nthreads < 1) nthreads = 1;
  mesh_job *jobs = (mesh_job *)calloc((size_t)nthreads, sizeof(mesh_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  for (int32_t i = 0; i < nthreads; i++) {
    jobs[i] = (mesh_job){(const uint8_t *)vox, es, greedy,
                         (int32_t)((int64_t)n_chunks * i / nthreads),
                         (int32_t)((int64_t)n_chunks * (i + 1) / nthreads), counts, 0};
    pthread_create(&th[i], NULL, mesh_worker, &jobs[i]);
  }
  uint64_t total = 0;
  for (int32_t i = 0; i < nthreads; i++) {
    pthread_join(th[i], NULL);
    total += jobs[i].total;
  }
  free(jobs);
  free(th);
  return total;
}

/* ------------------------------------------------------------------ */
/* reference restatements                                               */
/* ------------------------------------------------------------------ */

/* Erosion mip chain.  mip 0 is the occupancy.  mip d is set at a solid cell
 * iff the cell is at least 1 from every border (bIsOutOfBoundThickness,
 * VoxelMathHelper.h:106-112) and its 26 Moore neighbours (Get26Offsets,
 * BinaryOccupancyVolume.h:45-61) are all set in mip d-1; only solid cells are
 * visited (Chunk.h:86-91), all other cells stay 0. */
void mo_ref_erode_mips(const uint8_t *occ, int32_t res, int32_t depth, uint8_t *mips) {
  int64_t vol = (int64_t)res * res * res;
  memcpy(mips, occ, (size_t)vol);
  for (int32_t d = 1; d < depth; d++) {
    const uint8_t *prev = mips + (d - 1) * vol;
    uint8_t *cur = mips + d * vol;
    memset(cur, 0, (size_t)vol);
    for (int32_t z = 0; z < res; z++)
      for (int32_t y = 0; y < res; y++)
        for (int32_t x = 0; x < res; x++) {
          if (!occ[lin3(x, y, z, res)]) continue;
          if (x < 1 || y < 1 || z < 1 || x >= res - 1 || y >= res - 1 || z >= res - 1) continue;
          int all = 1;
          for (int dz = -1; dz <= 1 && all; dz++)
            for (int dy = -1; dy <= 1 && all; dy++)
              for (int dx = -1; dx <= 1 && all; dx++) {
                if (!dx && !dy && !dz) continue;
                if (!prev[lin3(x + dx, y + dy, z + dz, res)]) all = 0;
              }
          cur[lin3(x, y, z, res)] = (uint8_t)all;
        }
  }
}

/* GenerateSphere: FP64 signed distance at the cell corner
 * p = chunk*res*size + (X,Y,Z)*size to the sphere of radius 50 at (100,0,0);
 * cells emitted with X outermost and Z innermost. */
int32_t mo_ref_generate_sphere(int32_t cx, int32_t cy, int32_t cz, float block_size, int32_t res,
                               uint8_t *cells) {
  int32_t k = 0;
  double bs = (double)block_size, cr = (double)res;
  for (int32_t X = 0; X < res; X++)
    for (int32_t Y = 0; Y < res; Y++)
      for (int32_t Z = 0; Z < res; Z++) {
        double px = (double)cx * bs * cr + (double)X * bs;
        double py = (double)cy * bs * cr + (double)Y * bs;
        double pz = (double)cz * bs * cr + (double)Z * bs;
        double dx = px - 100.0, dy = py, dz = pz;
        double dd = dx * dx + dy * dy;
        dd = dd + dz * dz;
        double d = sqrt(dd) - 50.0;
        if (d < 0.0) {
          if (cells) {
            cells[3 * k + 0] = (uint8_t)X;
            cells[3 * k + 1] = (uint8_t)Y;
            cells[3 * k + 2] = (uint8_t)Z;
          }
          k++;
        }
      }
  return k;
}

static double ref_fract(double v) { return v - floor(v); }
static double ref_hash3(double x, double y, double z) {
  double dd = x * 127.1 + y * 311.7;
  dd = dd + z * 74.7;
  return ref_fract(sin(dd) * 43758.5453123);
}

/* gradient value noise with analytic derivatives; out = (dx, dy, dz, value) */
static void ref_noised(const double x[3], double out[4]) {
  double p[3], w[3], u[3], du[3];
  for (int i = 0; i < 3; i++) {
    p[i] = floor(x[i]);
    w[i] = x[i] - p[i];
    u[i] = w[i] * w[i] * w[i] * (w[i] * (w[i] * 6.0 - 15.0) + 10.0);
    du[i] = 30.0 * w[i] * w[i] * (w[i] * (w[i] - 2.0) + 1.0);
  }
  double a = ref_hash3(p[0], p[1], p[2]);
  double b = ref_hash3(p[0] + 1, p[1], p[2]);
  double c = ref_hash3(p[0], p[1] + 1, p[2]);
  double d = ref_hash3(p[0] + 1, p[1] + 1, p[2]);
  double e = ref_hash3(p[0], p[1], p[2] + 1);
  double f = ref_hash3(p[0] + 1, p[1], p[2] + 1);
  double g = ref_hash3(p[0], p[1] + 1, p[2] + 1);
  double h = ref_hash3(p[0] + 1, p[1] + 1, p[2] + 1);
  double k0 = a, k1 = b - a, k2 = c - a, k3 = e - a;
  double k4 = a - b - c + d, k5 = a - c - e + g, k6 = a - b - e + f;
  double k7 = -a + b + c - d + e - f - g + h;
  out[3] = -1.0 + 2.0 * (k0 + k1 * u[0] + k2 * u[1] + k3 * u[2] + k4 * u[0] * u[1] +
                         k5 * u[1] * u[2] + k6 * u[2] * u[0] + k7 * u[0] * u[1] * u[2]);
  out[0] = 2.0 * du[0] * (k1 + k4 * u[1] + k6 * u[2] + k7 * u[1] * u[2]);
  out[1] = 2.0 * du[1] * (k2 + k5 * u[2] + k4 * u[0] + k7 * u[2] * u[0]);
  out[2] = 2.0 * du[2] * (k3 + k6 * u[0] + k5 * u[1] + k7 * u[0] * u[1]);
}

static double ref_displacement(const double p0[3]) {
  double p[3] = {p0[0], p0[1], p0[2]};
  double mgn = 0.5, d = 0.0, s = 1.0, r[4];
  for (int i = 0; i < 5; i++) {
    double q[3] = {p[0] + 10.0, p[1] + 10.0, p[2] + 10.0};
    ref_noised(q, r);
    d += r[3] * mgn;
    for (int k = 0; k < 3; k++) p[k] = p[k] * 2.0 + r[k] * 0.2 * s;
    if (i == 2) s = -s;
    mgn *= 0.5;
  }
  for (int k = 0; k < 3; k++) p[k] = p0[k] * 32.0;
  for (int i = 0; i < 4; i++) {
    ref_noised(p, r);
    d += r[3] * mgn;
    for (int k = 0; k < 3; k++) p[k] *= 2.0;
    mgn *= 0.5;
  }
  return d;
}

double mo_ref_hash3(double x, double y, double z) { return ref_hash3(x, y, z); }
void mo_ref_noised(const double x[3], double out[4]) { ref_noised(x, out); }
double mo_ref_displacement(const double p[3]) { return ref_displacement(p); }

int32_t mo_ref_test_generator(int32_t cx, int32_t cy, int32_t cz, float block_size, int32_t res,
                              uint8_t *cells) {
  int32_t k = 0;
  double bs = (double)block_size, cr = (double)res;
  for (int32_t X = 0; X < res; X++)
    for (int32_t Y = 0; Y < res; Y++)
      for (int32_t Z = 0; Z < res; Z++) {
        double p[3] = {(double)cx * bs * cr + (double)X * bs, (double)cy * bs * cr + (double)Y * bs,
                       (double)cz * bs * cr + (double)Z * bs};
        double q[3] = {p[0] * 0.1, p[1] * 0.1, p[2] * 0.1};
        double d = (p[1] * 0.5 + ref_displacement(q) * 10.3) * 0.4;
        if (d < 0.0) {
          if (cells) {
            cells[3 * k + 0] = (uint8_t)X;
            cells[3 * k + 1] = (uint8_t)Y;
            cells[3 * k + 2] = (uint8_t)Z;
          }
          k++;
        }
      }
  return k;
}

int32_t mo_ref_octant_id(float x, float y, float z) {
  int32_t id = 0;
  if (x < 0.0f) id |= 1;
  if (y < 0.0f) id |= 2;
  if (z < 0.0f) id |= 4;
  return id;
}

/* ------------------------------------------------------------------ */
/* edits and scene chunks                                               */
/* ------------------------------------------------------------------ */

/* slot of the brick (bx, by, bz) of a brick map or two-level scene */
static uint32_t brick_slot(const mo_scene *s, int32_t bx, int32_t by, int32_t bz) {
  const int32_t nb = s->n / 8;
  if (bx < 0 || by < 0 || bz < 0 || bx >= nb || by >= nb || bz >= nb) return MO_EMPTY_SLOT;
  if (s->kind == MO_SCENE_BRICKMAP_U16) return s->slots[lin3(bx, by, bz, nb)];
  const uint32_t node = s->top[lin3(bx >> 4, by >> 4, bz >> 4, s->n / 128)];
  if (node == MO_EMPTY_SLOT) return MO_EMPTY_SLOT;
  return s->slots[(int64_t)node * 4096 + lin3(bx & 15, by & 15, bz & 15, 16)];
}

void mo_edit_chunks(const mo_scene *s, uint16_t *voxels, uint64_t *masks, const int32_t *xyz,
                    int32_t n_chunks, uint32_t seed) {
  const int32_t nc = (s->n + 31) / 32; /* a partial last chunk counts */
  for (int32_t k = 0; k < n_chunks; k++) {
    const int32_t cx = xyz[3 * k], cy = xyz[3 * k + 1], cz = xyz[3 * k + 2];
    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) continue;
    const uint32_t h0 = mo_hash32(seed ^ mo_hash32((uint32_t)lin3(cx, cy, cz, nc)));
    const uint32_t h1 = mo_hash32(h0), h2 = mo_hash32(h1), h3 = mo_hash32(h2);
    const int64_t sx = (int64_t)cx * 8192 + (h0 & 0x1FFF), sy = (int64_t)cy * 8192 + (h1 & 0x1FFF),
                  sz = (int64_t)cz * 8192 + (h2 & 0x1FFF);
    const int64_t r = 4 * 256 + (h3 % (8u * 256u));
    for (int32_t b = 0; b < 64; b++) {
      const int32_t bx = cx * 4 + (b & 3), by = cy * 4 + ((b >> 2) & 3), bz = cz * 4 + (b >> 4);
      const uint32_t slot = brick_slot(s, bx, by, bz);
      if (slot == MO_EMPTY_SLOT) continue;
      for (int32_t i = 0; i < 512; i++) {
        const int32_t lx = i & 7, ly = (i >> 3) & 7, lz = i >> 6;
        const int64_t dx = (int64_t)(bx * 8 + lx) * 256 + 128 - sx;
        const int64_t dy = (int64_t)(by * 8 + ly) * 256 + 128 - sy;
        const int64_t dz = (int64_t)(bz * 8 + lz) * 256 + 128 - sz;
        if (dx * dx + dy * dy + dz * dz > r * r) continue;
        uint16_t *v = voxels + (int64_t)slot * 512 + i;
        *v = *v ? 0 : 7;
      }
      for (int32_t lz = 0; lz < 8; lz++) {
        uint64_t m = 0;
        for (int32_t j = 0; j < 64; j++)
          if (voxels[(int64_t)slot * 512 + lz * 64 + j]) m |= 1ull << j;
        masks[(int64_t)slot * 8 + lz] = m;
      }
    }
  }
}

void mo_scene_box(const mo_scene *s, int32_t lo[3], int32_t hi[3]) {
  if (s->kind == MO_SCENE_DENSE_U8) {
    for (int a = 0; a < 3; a++) {
      lo[a] = 0;
      hi[a] = s->n;
    }
    return;
  }
  const int32_t nb = s->n / 8;
  int32_t mn[3] = {nb, nb, nb}, mx[3] = {-1, -1, -1};
  for (int32_t bz = 0; bz < nb; bz++)
    for (int32_t by = 0; by < nb; by++)
      for (int32_t bx = 0; bx < nb; bx++) {
        if (brick_slot(s, bx, by, bz) == MO_EMPTY_SLOT) continue;
        const int32_t b[3] = {bx, by, bz};
        for (int a = 0; a < 3; a++) {
          if (b[a] < mn[a]) mn[a] = b[a];
          if (b[a] > mx[a]) mx[a] = b[a];
        }
      }
  for (int a = 0; a < 3; a++) {
    lo[a] = mx[a] < 0 ? 0 : mn[a] * 8;
    hi[a] = mx[a] < 0 ? 0 : (mx[a] + 1) * 8;
  }
}

void mo_scene_chunk(const mo_scene *s, int32_t cx, int32_t cy, int32_t cz, uint16_t *out) {
  for (int32_t z = 0; z < 32; z++)
    for (int32_t y = 0; y < 32; y++)
      for (int32_t x = 0; x < 32; x++) {
        const int32_t gx = cx * 32 + x, gy = cy * 32 + y, gz = cz * 32 + z;
        const uint32_t slot = brick_slot(s, gx >> 3, gy >> 3, gz >> 3);
        out[x + 32 * y + 1024 * z] =
            slot == MO_EMPTY_SLOT ? 0
                                  : s->voxels[(int64_t)slot * 512 + (gx & 7) + 8 * (gy & 7) + 64 * (gz & 7)];
      }
}

/* ------------------------------------------------------------------ */
/* chunk pool                                                           */
/* ------------------------------------------------------------------ */

float mo_pool_view_threshold(float view_angle_deg) {
  float r = view_angle_deg * 0.017453292f; /* degrees to radians */
  r = r * 0.5f;
  float c = cosf(r);
  return c > 0.01f ? c : 0.01f;
}

void mo_pool_forward(const float raw[3], float f[3]) {
  float ss = raw[0] * raw[0] + raw[1] * raw[1];
  ss = ss + raw[2] * raw[2];
  const float len = sqrtf(ss);
  for (int i = 0; i < 3; i++) f[i] = raw[i] / len;
}

/* GetFibonacciSphere<float> (Runtimes/Helper/VoxelMathHelper.h:47-68), as
 * BakeVisibilityByView instantiates it (ChunkManagerHelper.h:205): Phi is
 * computed in double and stored as float; everything after is float
 * arithmetic (cosf / sinf / sqrtf), and glm::normalize multiplies by
 * 1 / sqrt(dot(v, v)) with dot summed x, y, then z. */
void mo_pool_bake_directions(int32_t n, float *dirs) {
  const float phi = (float)(3.14159265358979323846 * (sqrt(5.0) - 1.0));
  for (int32_t i = 0; i < n; i++) {
    float y = (float)i / (float)(n - 1);
    y = y * 2.0f;
    y = 1.0f - y;
    float yy = y * y;
    const float r = sqrtf(1.0f - yy);
    const float th = phi * (float)i;
    const float x = cosf(th) * r;
    const float z = sinf(th) * r;
    float dd = x * x + yy;
    dd = dd + z * z;
    const float inv = 1.0f / sqrtf(dd);
    dirs[3 * i + 0] = x * inv;
    dirs[3 * i + 1] = y * inv;
    dirs[3 * i + 2] = z * inv;
  }
}

void mo_pool_snap(const float *dirs, int32_t n, const float f[3], float out[3]) {
  int32_t best = 0;
  float bd = INFINITY;
  for (int32_t i = 0; i < n; i++) {
    const float dx = dirs[3 * i] - f[0], dy = dirs[3 * i + 1] - f[1], dz = dirs[3 * i + 2] - f[2];
    float d = dx * dx + dy * dy;
    d = d + dz * dz;
    if (d < bd) {
      bd = d;
      best = i;
    }
  }
  for (int k = 0; k < 3; k++) out[k] = dirs[3 * best + k];
}

/* length of an integer offset and the dot of its direction with f */
static float offset_dir_dot(int32_t dx, int32_t dy, int32_t dz, const float f[3], float *len) {
  const float x = (float)dx, y = (float)dy, z = (float)dz;
  float ss = x * x + y * y;
  ss = ss + z * z;
  *len = sqrtf(ss);
  const float ux = x / *len, uy = y / *len, uz = z / *len;
  float d = ux * f[0] + uy * f[1];
  d = d + uz * f[2];
  return d;
}

float mo_chunk_importance(int32_t dx, int32_t dy, int32_t dz, const float f[3]) {
  if (dx >= -2 && dx <= 2 && dy >= -2 && dy <= 2 && dz >= -2 && dz <= 2) return 1.0e6f;
  float len;
  float a = offset_dir_dot(dx, dy, dz, f, &len);
  a = a > 0.0f ? a : 0.0f;
  a = a - 0.5f;
  a = a * 2.0f;
  a = a > 0.75f ? a : 0.75f;
  float di = 64.0f - len;
  di = di > 0.25f ? di : 0.25f;
  return a * di;
}

int mo_chunk_desired(int32_t dx, int32_t dy, int32_t dz, const float f[3],
                     const mo_pool_params *p) {
  const float x = (float)dx, y = (float)dy, z = (float)dz;
  float ss = x * x + y * y;
  ss = ss + z * z;
  const float len0 = sqrtf(ss);
  if ((double)len0 > (double)p->forward_radius + 1e-6) return 0;
  if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) return 1;
  float len;
  float a = offset_dir_dot(dx, dy, dz, f, &len);
  a = a > 0.0f ? a : 0.0f;
  a = a > p->view_threshold ? 1.0f : a / p->view_threshold;
  a = a < 0.0f ? 0.0f : (a > 1.0f ? 1.0f : a);
  const float t1 = a * (float)p->forward_radius;
  const float b = 1.0f - a;
  const float t2 = b * (float)p->backward_radius;
  const float thr = t1 + t2;
  return len < thr;
}

int mo_chunk_nonempty(const mo_scene *s, int32_t cx, int32_t cy, int32_t cz) {
  const int32_t nc = (s->n + 31) / 32;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) return 0;
  if (s->kind == MO_SCENE_BRICKMAP_U16) return get_bit(s->occ32, lin3(cx, cy, cz, nc));
  const uint32_t node = s->top[lin3(cx >> 2, cy >> 2, cz >> 2, s->n / 128)];
  if (node == MO_EMPTY_SLOT) return 0;
  return (int)((s->groups[node] >> ((cx & 3) + 4 * (cy & 3) + 16 * (cz & 3))) & 1u);
}

static uint32_t float_bits(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return u;
}

static int cmp_u64(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

void mo_pool_update(const mo_scene *s, const mo_pool_params *p, uint32_t stamp,
                    const int32_t cam[3], const float f[3], uint32_t *lookup,
                    uint32_t *slot_chunk, uint32_t *slot_stamp, uint32_t *slot_count,
                    uint32_t *quads, int32_t *load_xyz, uint32_t *load_slot,
                    mo_pool_result *res) {
  const int32_t nc = (s->n + 31) / 32, R = p->forward_radius, side = 2 * R + 1;
  const int64_t cube = (int64_t)side * side * side;
  uint64_t *cand = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)cube);
  uint64_t *vict = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)p->capacity);
  uint32_t n_cand = 0;
  for (int64_t k = 0; k < cube; k++) {
    const int32_t dx = (int32_t)(k % side) - R, dy = (int32_t)((k / side) % side) - R,
                  dz = (int32_t)(k / ((int64_t)side * side)) - R;
    const int32_t cx = cam[0] + dx, cy = cam[1] + dy, cz = cam[2] + dz;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc) continue;
    const uint32_t idx = (uint32_t)lin3(cx, cy, cz, nc);
    if (lookup[idx] != MO_NO_SLOT || !mo_chunk_desired(dx, dy, dz, f, p) ||
        !mo_chunk_nonempty(s, cx, cy, cz))
      continue;
    const float imp = mo_chunk_importance(dx, dy, dz, f);
    cand[n_cand++] = ((uint64_t)(~float_bits(imp)) << 32) | idx;
  }
  qsort(cand, n_cand, sizeof(uint64_t), cmp_u64);
  for (int32_t sl = 0; sl < p->capacity; sl++) {
    const uint32_t c = slot_chunk[sl];
    if (c == MO_NO_SLOT) {
      vict[sl] = (uint64_t)sl;
      continue;
    }
    const int32_t x = (int32_t)(c % (uint32_t)nc), y = (int32_t)((c / (uint32_t)nc) % (uint32_t)nc),
                  z = (int32_t)(c / ((uint32_t)nc * (uint32_t)nc));
    const float imp = mo_chunk_importance(x - cam[0], y - cam[1], z - cam[2], f);
    vict[sl] = ((uint64_t)float_bits(imp) << 32) | (uint32_t)sl;
  }
  qsort(vict, (size_t)p->capacity, sizeof(uint64_t), cmp_u64);
  uint32_t n = 0, evicted = 0;
  while (n < n_cand && (int32_t)n < p->max_loads && (int32_t)n < p->capacity) {
    const uint32_t vb = (uint32_t)(vict[n] >> 32);
    float vi, ci;
    const uint32_t cb = ~(uint32_t)(cand[n] >> 32);
    memcpy(&vi, &vb, 4);
    memcpy(&ci, &cb, 4);
    if (vb != 0 && !(vi < ci)) break;
    n++;
  }
  uint16_t *vox = (uint16_t *)malloc(32768 * sizeof(uint16_t));
  uint32_t *tmp = (uint32_t *)malloc(98304 * sizeof(uint32_t));
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t sl = (uint32_t)vict[i], c = (uint32_t)cand[i];
    const uint32_t old = slot_chunk[sl];
    if (old != MO_NO_SLOT) {
      lookup[old] = MO_NO_SLOT;
      evicted++;
    }
    lookup[c] = sl;
    slot_chunk[sl] = c;
    slot_stamp[sl] = stamp;
    const int32_t x = (int32_t)(c % (uint32_t)nc), y = (int32_t)((c / (uint32_t)nc) % (uint32_t)nc),
                  z = (int32_t)(c / ((uint32_t)nc * (uint32_t)nc));
    load_xyz[3 * i] = x;
    load_xyz[3 * i + 1] = y;
    load_xyz[3 * i + 2] = z;
    load_slot[i] = sl;
    mo_scene_chunk(s, x, y, z, vox);
    const uint32_t cnt = mo_mesh_chunk(vox, 2, p->greedy, tmp, 98304);
    const uint32_t keep = cnt < (uint32_t)p->quads_per_slot ? cnt : (uint32_t)p->quads_per_slot;
    memcpy(quads + (uint64_t)sl * (uint32_t)p->quads_per_slot, tmp, keep * sizeof(uint32_t));
    slot_count[sl] = cnt;
  }
  free(tmp);
  free(vox);
  free(cand);
  free(vict);
  res->n_candidates = n_cand;
  res->n_loads = n;
  res->n_evicted = evicted;
}

/* ------------------------------------------------------------------ */
/* quad visibility                                                      */
/* ------------------------------------------------------------------ */

typedef struct {
  int a, u, v;       /* plane axis and the two in-plane axes */
  int pos;           /* 1: the face normal points along +a */
  float P;           /* plane coordinate */
  int32_t U0, V0, W, H;
  int32_t slice;     /* voxel coordinate along a of the face's voxel */
} quad_geom;

static void quad_decode(uint32_t q, int32_t ox, int32_t oy, int32_t oz, quad_geom *g) {
  const int32_t x = (int32_t)(q & 31u), y = (int32_t)((q >> 5) & 31u), z = (int32_t)((q >> 10) & 31u);
  const int32_t w = (int32_t)((q >> 15) & 31u) + 1, h = (int32_t)((q >> 20) & 31u) + 1;
  const int f = (int)((q >> 25) & 7u);
  const int32_t X = ox + x, Y = oy + y, Z = oz + z;
  g->pos = (f & 1) == 0;
  if (f < 2) {
    g->a = 2; g->u = 0; g->v = 1; g->slice = Z; g->U0 = X; g->V0 = Y;
  } else if (f < 4) {
    g->a = 1; g->u = 0; g->v = 2; g->slice = Y; g->U0 = X; g->V0 = Z;
  } else {
    g->a = 0; g->u = 1; g->v = 2; g->slice = X; g->U0 = Y; g->V0 = Z;
  }
  g->W = w;
  g->H = h;
  g->P = (float)(g->slice + (g->pos ? 1 : 0));
}

/* t of the ray's hit on the quad, or -1 */
static float quad_hit(const quad_geom *g, const float o[3], const float d[3]) {
  const float da = d[g->a];
  if (g->pos ? !(da < 0.0f) : !(da > 0.0f)) return -1.0f;
  const float inv = 1.0f / da;
  const float q = g->P - o[g->a];
  const float t = q * inv;
  if (!(t > 0.0f)) return -1.0f;
  const float du = d[g->u] * t, dv = d[g->v] * t;
  const float hu = o[g->u] + du, hv = o[g->v] + dv;
  if (!(hu >= (float)g->U0 && hu < (float)(g->U0 + g->W))) return -1.0f;
  if (!(hv >= (float)g->V0 && hv < (float)(g->V0 + g->H))) return -1.0f;
  return t;
}

/* conservative pixel rectangle [x0, x1) x [y0, y1) of a quad; 0 if empty */
static int quad_rect(const mo_camera *c, const quad_geom *g, int32_t r[4]) {
  double cam[4][3], poly[8][3];
  for (int k = 0; k < 4; k++) {
    double p[3];
    p[g->a] = g->P;
    p[g->u] = g->U0 + ((k == 1 || k == 2) ? g->W : 0);
    p[g->v] = g->V0 + ((k >= 2) ? g->H : 0);
    double e[3] = {p[0] - c->eye[0], p[1] - c->eye[1], p[2] - c->eye[2]};
    cam[k][0] = e[0] * c->right[0] + e[1] * c->right[1] + e[2] * c->right[2];
    cam[k][1] = e[0] * c->up[0] + e[1] * c->up[1] + e[2] * c->up[2];
    cam[k][2] = e[0] * c->fwd[0] + e[1] * c->fwd[1] + e[2] * c->fwd[2];
  }
  const double zn = 1e-3;
  int m = 0;
  for (int k = 0; k < 4; k++) { /* clip the polygon at z >= zn */
    const double *A = cam[k], *B = cam[(k + 1) & 3];
    const int ia = A[2] >= zn, ib = B[2] >= zn;
    if (ia) {
      poly[m][0] = A[0]; poly[m][1] = A[1]; poly[m][2] = A[2]; m++;
    }
    if (ia != ib) {
      const double s = (zn - A[2]) / (B[2] - A[2]);
      for (int i = 0; i < 3; i++) poly[m][i] = A[i] + (B[i] - A[i]) * s;
      m++;
    }
  }
  if (m == 0) return 0;
  double x0 = 1e30, x1 = -1e30, y0 = 1e30, y1 = -1e30;
  for (int k = 0; k < m; k++) {
    const double ax = poly[k][0] / poly[k][2], by = poly[k][1] / poly[k][2];
    const double px = (ax / c->sx + 1.0) * 0.5 * c->width - 0.5;
    const double py = (1.0 - by / c->sy) * 0.5 * c->height - 0.5;
    x0 = px < x0 ? px : x0; x1 = px > x1 ? px : x1;
    y0 = py < y0 ? py : y0; y1 = py > y1 ? py : y1;
  }
  double fx0 = floor(x0) - 1.0, fx1 = ceil(x1) + 2.0, fy0 = floor(y0) - 1.0, fy1 = ceil(y1) + 2.0;
  if (fx0 < 0) fx0 = 0;
  if (fy0 < 0) fy0 = 0;
  if (fx1 > c->width) fx1 = c->width;
  if (fy1 > c->height) fy1 = c->height;
  if (fx0 >= fx1 || fy0 >= fy1) return 0;
  r[0] = (int32_t)fx0; r[1] = (int32_t)fx1; r[2] = (int32_t)fy0; r[3] = (int32_t)fy1;
  return 1;
}

void mo_raster_pool(const mo_camera *cam, int32_t n, int32_t nc, const uint32_t *slot_chunk,
                    const uint32_t *slot_count, const uint32_t *quads, int32_t capacity,
                    int32_t per_slot, const float light[3], uint64_t *keys, int64_t *out_index,
                    uint8_t *out_face, float *out_t, float *out_shade) {
  const int64_t px_n = (int64_t)cam->width * cam->height;
  for (int64_t i = 0; i < px_n; i++) keys[i] = ~0ull;
  for (int32_t sl = 0; sl < capacity; sl++) {
    const uint32_t ch = slot_chunk[sl];
    if (ch == MO_NO_SLOT) continue;
    const int32_t ox = (int32_t)(ch % (uint32_t)nc) * 32, oy = (int32_t)((ch / (uint32_t)nc) % (uint32_t)nc) * 32,
                  oz = (int32_t)(ch / ((uint32_t)nc * (uint32_t)nc)) * 32;
    const uint32_t cnt = slot_count[sl] < (uint32_t)per_slot ? slot_count[sl] : (uint32_t)per_slot;
    for (uint32_t k = 0; k < cnt; k++) {
      quad_geom g;
      quad_decode(quads[(uint64_t)sl * (uint32_t)per_slot + k], ox, oy, oz, &g);
      if (g.pos ? !(cam->eye[g.a] > g.P) : !(cam->eye[g.a] < g.P)) continue; /* back-facing */
      int32_t r[4];
      if (!quad_rect(cam, &g, r)) continue;
      const uint64_t id = (uint64_t)sl * (uint32_t)per_slot + k;
      for (int32_t py = r[2]; py < r[3]; py++)
        for (int32_t px = r[0]; px < r[1]; px++) {
          float o[3], d[3];
          mo_primary_ray(cam, px, py, o, d);
          const float t = quad_hit(&g, o, d);
          if (t < 0.0f) continue;
          const uint64_t key = ((uint64_t)float_bits(t) << 32) | id;
          uint64_t *kp = &keys[(int64_t)py * cam->width + px];
          if (key < *kp) *kp = key;
        }
    }
  }
  for (int64_t i = 0; i < px_n; i++) {
    const uint64_t key = keys[i];
    if (key == ~0ull) {
      out_index[i] = -1;
      out_face[i] = MO_FACE_MISS;
      out_t[i] = INFINITY;
      out_shade[i] = 0.0f;
      continue;
    }
    const uint32_t id = (uint32_t)key, sl = id / (uint32_t)per_slot;
    const uint32_t ch = slot_chunk[sl], q = quads[id];
    quad_geom g;
    quad_decode(q, (int32_t)(ch % (uint32_t)nc) * 32, (int32_t)((ch / (uint32_t)nc) % (uint32_t)nc) * 32,
                (int32_t)(ch / ((uint32_t)nc * (uint32_t)nc)) * 32, &g);
    const int32_t px = (int32_t)(i % cam->width), py = (int32_t)(i / cam->width);
    float o[3], d[3];
    mo_primary_ray(cam, px, py, o, d);
    const uint32_t tb = (uint32_t)(key >> 32);
    float t;
    memcpy(&t, &tb, 4);
    const float du = d[g.u] * t, dv = d[g.v] * t;
    const float hu = o[g.u] + du, hv = o[g.v] + dv;
    int32_t cu = (int32_t)floorf(hu), cv = (int32_t)floorf(hv);
    cu = cu < g.U0 ? g.U0 : (cu > g.U0 + g.W - 1 ? g.U0 + g.W - 1 : cu);
    cv = cv < g.V0 ? g.V0 : (cv > g.V0 + g.H - 1 ? g.V0 + g.H - 1 : cv);
    int32_t cell[3];
    cell[g.a] = g.slice;
    cell[g.u] = cu;
    cell[g.v] = cv;
    const int f = (int)((q >> 25) & 7u);
    out_index[i] = lin3(cell[0], cell[1], cell[2], n);
    out_face[i] = (uint8_t)f;
    out_t[i] = t;
    const int sg = (f & 1) ? -1 : 1;
    float lambert = sg > 0 ? light[g.a] : -light[g.a];
    if (lambert < 0.0f) lambert = 0.0f;
    float l = lambert * 1.0f;
    l = 0.75f * l;
    float shade = 0.25f + l;
    shade = shade * 1.0f;
    out_shade[i] = shade;
  }
}
