// Device-side voxel traversal for gfx950: Amanatides & Woo (1987) DDA with
// hierarchical empty-space skipping over 128^3 blocks, 32^3 blocks, 8^3
// bricks, the 4^3 and 2^3 blocks inside a brick, and voxels.  Every float operation is written out in a fixed order
// and the library is compiled with -ffp-contract=off, so results are the IEEE
// binary32 results of these operations: the integer outputs (voxel index,
// face) and the float outputs (t, shade) are reproducible bit for bit by any
// scalar IEEE implementation of the same steps.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "meso_internal.h"

namespace meso {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kFaceInside = 6;
constexpr int kFaceMiss = 0xFF;

// Face ids follow Runtimes/Shape/Cube.h:25-54: 0:+Z 1:-Z 2:+Y 3:-Y 4:+X 5:-X.
// A ray stepping +x enters a voxel through its -X face.
__device__ __forceinline__ int entry_face(int axis, int step) {
  int base = axis == 2 ? 0 : (axis == 1 ? 2 : 4);
  return base + (step > 0 ? 1 : 0);
}
__device__ __forceinline__ int face_axis(int f) { return f < 2 ? 2 : (f < 4 ? 1 : 0); }

struct Ray {
  float o[3], d[3], inv[3];  // inv: +inf where d = 0
  int32_t step[3];
  // the exit plane of block b is b + pos: 1 where step > 0, 0 where step < 0,
  // and 2 where step = 0, so that plane lies beyond the origin and its t,
  // (plane - o) * inf, is +inf: an axis the ray does not move along is never
  // the nearest, with no select per step
  int32_t pos[3];
  uint32_t posmask;  // bit a: step[a] > 0
  uint32_t negmask;  // bit a: step[a] < 0 (the next cell is one below the plane)
  // the direction the cell can move in on axis a: step[a] > 0, or on an
  // axis the ray does not move along, floor(o[a]) above the start cell
  bool up[3];
};

__device__ __forceinline__ void ray_setup(Ray &r, const float o[3], const float d[3]) {
  r.posmask = 0u;
  r.negmask = 0u;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    r.o[a] = o[a];
    r.d[a] = d[a];
    // one division per axis, kept where d != 0: written per branch, the
    // divide sequence is emitted twice and lanes of both signs run both
    const float inv = 1.0f / d[a];
    const bool pos = d[a] > 0.0f, neg = d[a] < 0.0f;
    r.step[a] = pos ? 1 : (neg ? -1 : 0);
    r.inv[a] = (pos || neg) ? inv : __builtin_inff();
    r.pos[a] = pos ? 1 : (neg ? 0 : 2);
    r.up[a] = pos;
    r.posmask |= (r.step[a] > 0 ? 1u : 0u) << a;
    r.negmask |= (r.step[a] < 0 ? 1u : 0u) << a;
  }
}

__device__ __forceinline__ float plane_t(float o, float inv, float b) {
  float q = b - o;
  return q * inv;
}
// smallest t, ties to x then y
__device__ __forceinline__ int argmin3(float t0, float t1, float t2) {
  if (t0 <= t1 && t0 <= t2) return 0;
  return t1 <= t2 ? 1 : 2;
}

struct WalkState {
  int32_t cell[3];
  float t;
  int32_t face;
};

struct StepCounts {
  uint32_t n_vox, n_top, n_node, n_brickslot, n_brickmask, n_coarse, n_iter;
};

__device__ __forceinline__ int64_t lin3(int64_t x, int64_t y, int64_t z, int64_t n) {
  return x + y * n + z * n * n;
}
// Index of (x,y,z) in an n^3 grid of blocks or bricks, with 24-bit multiplies
// (full rate; a 32-bit multiply is quarter rate): scene edges are at most
// 8192, so n <= 1024 and x, y, z < n, and every operand stays below 2^24 and
// every product below 2^32.
__device__ __forceinline__ uint32_t lin3u(uint32_t x, uint32_t y, uint32_t z, uint32_t n) {
  return x + __umul24(y + __umul24(z, n), n);
}

// The median of a, b, c: for lo <= hi, med3_i32(x, lo, hi) clamps x into
// [lo, hi] in one instruction (the compiler cannot see lo <= hi).
__device__ __forceinline__ int32_t med3_i32(int32_t a, int32_t b, int32_t c) {
  int32_t m;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(m) : "v"(a), "v"(b), "v"(c));
  return m;
}

// (int32_t)floor(x) in one instruction: on gfx950, v_cvt_flr_i32_f32 gives
// the same result as v_floor_f32 then v_cvt_i32_f32 for every finite float
// (checked on the device over every finite float by tests/native/floor_convert_probe.hip)
__device__ __forceinline__ int32_t floor_i32(float x) {
  int32_t r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// (1 << k) - 1 for k in [0, 31], one instruction
__device__ __forceinline__ int32_t low_mask(int32_t k) {
  int32_t m;
  asm("v_bfm_b32 %0, %1, 0" : "=v"(m) : "v"(k));
  return m;
}

// One step at level 2^sh (sh = 0 voxel, 3 brick, 5 and 7 empty blocks): leave
// the block holding the current cell through its nearest plane.  The other
// two coordinates come from the exit point, clamped into the block and the
// scene and never moving against the ray; at sh = 0 the block is the cell
// itself, so the clamp keeps them, which is the plain voxel step.  Written
// without branches (selects only), so that the lanes of a wave stepping at
// different levels stay converged.  Returns false when the step along the
// chosen axis leaves the scene's solid box (nothing solid lies beyond it);
// the step is computed all the same (no branch), and a walk that gets false
// ends without reading the cell.
template <bool LEVELS = true>
__device__ __forceinline__ bool advance(const Ray &r, const DevScene &s, int32_t sh, WalkState &w) {
  const int32_t n = s.n;
  float tb[3];
  int32_t bc[3], pl[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    bc[a] = w.cell[a] >> sh;
    pl[a] = (bc[a] + r.pos[a]) << sh;  // the exit plane on this axis
    // +inf where step = 0: pl is at least the cell + 2 there, and the
    // origin's coordinate stays below the cell + 2 (the cell is floor(o),
    // or for a secondary ray the hit cell, whose face holds o)
    tb[a] = plane_t(r.o[a], r.inv[a], (float)pl[a]);
  }
  // the chosen axis by selects and bit extracts only: an index into the
  // ray's arrays would move them out of registers
  const int ax = argmin3(tb[0], tb[1], tb[2]);
  const bool a0 = ax == 0, a1 = ax == 1;
  const float t = a0 ? tb[0] : (a1 ? tb[1] : tb[2]);
  const int32_t pax = a0 ? pl[0] : (a1 ? pl[1] : pl[2]);
  const int32_t na = pax - (int32_t)((r.negmask >> ax) & 1u);
  const int32_t face = 4 - 2 * ax + (int32_t)((r.posmask >> ax) & 1u);  // entry_face
  // compared with each axis' bounds (scalar operands; one unsigned compare
  // per axis) and the chosen axis' answer kept, combined as lane masks:
  // selecting the bounds per lane would hold them in vector registers for
  // the whole walk
  const bool out0 = (uint32_t)(na - s.lo[0]) >= (uint32_t)(s.hi[0] - s.lo[0]);
  const bool out1 = (uint32_t)(na - s.lo[1]) >= (uint32_t)(s.hi[1] - s.lo[1]);
  const bool out2 = (uint32_t)(na - s.lo[2]) >= (uint32_t)(s.hi[2] - s.lo[2]);
  const bool out = (a0 & out0) | (a1 & out1) | (!a0 & !a1 & out2);
  if constexpr (!LEVELS) {  // voxel steps only (dense grid): the other cells stay
#pragma unroll
    for (int b = 0; b < 3; b++) w.cell[b] = b == ax ? na : w.cell[b];
    w.t = t;
    w.face = face;
    return !out;
  }
  const int32_t span = low_mask(sh);
#pragma unroll
  for (int b = 0; b < 3; b++) {
    float dt = r.d[b] * t;
    float p = r.o[b] + dt;
    int32_t c = floor_i32(p);
    // clamp into the block [lo, hi], then not below the cell when stepping
    // up or above it when stepping down: the cell lies in [lo, hi], so that
    // is one clamp between the cell and one bound (a single med3).  On an
    // axis the ray does not move along, c is floor(o) on every step and
    // the cell only moves toward it, so the side up[] fixed at the start
    // holds for the whole walk and the same clamp gives clamp(c, lo, hi).
    const int32_t lo = bc[b] << sh;
    int32_t hi = lo + span;
    hi = hi > n - 1 ? n - 1 : hi;
    c = med3_i32(c, w.cell[b], r.up[b] ? hi : lo);
    w.cell[b] = b == ax ? na : c;
  }
  w.t = t;
  w.face = face;
  return !out;
}

// The 4x4 2^2 blocks of an 8x8 plane of a brick (bit x + 8 y) that hold a set
// bit, as 16 bits (bit cx + 4 cy): fold x and y pairs onto the even bits,
// then gather the 16 even bits.
__device__ __forceinline__ uint32_t plane_blocks2(uint64_t w) {
  w |= w >> 1;
  w |= w >> 8;
  w &= 0x0055005500550055ull;
  w = (w | (w >> 1)) & 0x0033003300330033ull;
  w = (w | (w >> 2)) & 0x000F000F000F000Full;
  w = (w | (w >> 12)) & 0x000000FF000000FFull;
  w = (w | (w >> 24)) & 0xFFFFull;
  return (uint32_t)w;
}

// One bit per aligned 2^3 block of a brick (bit cx + 4 cy + 16 cz), from its
// 64-byte mask (word z, bit x + 8 y): the planes of each z pair ORed, folded.
__device__ __forceinline__ uint64_t brick_blocks2(const uint64_t *mask) {
  uint64_t sub = 0;
#pragma unroll
  for (int p = 0; p < 4; p++)
    sub |= (uint64_t)plane_blocks2(mask[2 * p] | mask[2 * p + 1]) << (16 * p);
  return sub;
}

// A brick written by 512 lanes (lane = voxel x + 8 y + 64 z, each wave one z
// plane) whose plane bits are `b` (the wave's ballot of solid voxels): the
// brick's 2^3-block bits, returned to lane 0.  Every lane of the workgroup
// must call it (it synchronises the workgroup).
__device__ __forceinline__ uint64_t brick_blocks2_from_planes(uint64_t b) {
  __shared__ uint32_t plane[8];
  if ((threadIdx.x & 63u) == 0) plane[threadIdx.x >> 6] = plane_blocks2(b);
  __syncthreads();
  uint64_t sub = 0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int p = 0; p < 4; p++) sub |= (uint64_t)(plane[2 * p] | plane[2 * p + 1]) << (16 * p);
  }
  return sub;
}

__device__ __forceinline__ bool occ_bit(const uint32_t *bits, uint32_t i) {
  return (bits[i >> 5] >> (i & 31u)) & 1u;
}

// Cached answers of the coarse levels for the cell a walk is in.  The valid
// levels always form a prefix (the 128^3 block, then the 32^3 block, then
// the brick), and each valid level's block holds the walk's previous cell.
// So the cell is still in a valid level's block of edge 2^k exactly when dv,
// the previous step's (x ^ x') | (y ^ y') | (z ^ z'), is below 2^k, and
// max(dv, vmask) >= 2^k, with vmask one below the edge of the finest valid
// level (255: none valid), says that level k must be looked up.
struct WalkCache {
  uint32_t dv, vmask;
  bool e128, e32;
  uint32_t node, slot;
  uint64_t b32, sub;  // b32: allocated bricks of the 32^3 block; sub: 2^3-block bits of the brick
  __device__ __forceinline__ void reset() {
    dv = 0;
    vmask = 255u;
    e128 = e32 = false;
    node = slot = kEmpty;
    b32 = sub = 0;
  }
};

enum : int { kWalkOn = 0, kWalkHit = 1, kWalkOut = 2 };

// One step of a walk from a cell whose entry t is below tmax: kWalkHit when
// the cell is solid (material set), kWalkOut when the ray leaves the scene or
// the next cell's entry t reaches tmax, else kWalkOn after stepping to the
// next cell.  (Testing the next cell's t here, not at the top of the next
// step, saves a branch per step; walk tests the first cell's.)  KIND: 1 dense u8 (voxel
// steps only), 2 brick map, 3 two-level.  A brick scene steps at the coarsest
// empty level holding the cell: 128^3 block, 32^3 block (its word of
// allocated bricks, which also names the empty bricks without reading their
// slots), brick, the 4^3 and 2^3 blocks of a non-empty brick (from its 2^3-
// block bits, read once per brick), else the voxel (one 8-byte mask word).
template <int KIND, bool COUNT>
__device__ __forceinline__ int walk_step(const DevScene &s, const Ray &r, WalkState &w,
                                         WalkCache &c, float tmax, uint16_t &material,
                                         StepCounts &cnt) {
  const int32_t n = s.n;
  if (COUNT) cnt.n_iter++;
  if constexpr (KIND == 1) {
    if (COUNT) cnt.n_vox++;
    const uint8_t v = s.dense[lin3(w.cell[0], w.cell[1], w.cell[2], n)];
    if (v) {
      material = v;
      return kWalkHit;
    }
    return advance<false>(r, s, 0, w) && w.t < tmax ? kWalkOn : kWalkOut;
  } else {
    const uint32_t n128 = (uint32_t)(n + 127) >> 7, n32 = (uint32_t)(n + 31) >> 5;
    // one flat pass over the levels (no nesting, so a wave whose lanes sit
    // at different levels runs each piece once): refresh a level's cached
    // answer when the cell left its block, then step at the coarsest empty
    // level, or stop at a solid voxel
    const uint32_t x = (uint32_t)w.cell[0], y = (uint32_t)w.cell[1], z = (uint32_t)w.cell[2];
    const uint32_t q = c.dv > c.vmask ? c.dv : c.vmask;
    if (q >= 128u) {
      if constexpr (KIND == 3) {
        if (COUNT) cnt.n_top++;
        c.node = s.top[lin3u(x >> 7, y >> 7, z >> 7, n128)];
        c.e128 = c.node == kEmpty;
      } else {
        if (COUNT) cnt.n_coarse++;
        c.e128 = !occ_bit(s.occ128, lin3u(x >> 7, y >> 7, z >> 7, n128));
      }
    }
    if (!c.e128 && q >= 32u) {
      if (COUNT) cnt.n_coarse++;
      c.b32 = s.bricks32[lin3u(x >> 5, y >> 5, z >> 5, n32)];
      c.e32 = c.b32 == 0;
    }
    const bool coarse = c.e128 || c.e32;
    if (!coarse && q >= 8u) {
      if (COUNT) {
        if constexpr (KIND == 3)
          cnt.n_node++;
        else
          cnt.n_brickslot++;
      }
      c.slot = kEmpty;
      if ((c.b32 >> (((x >> 3) & 3u) + 4u * ((y >> 3) & 3u) + 16u * ((z >> 3) & 3u))) & 1ull) {
        // the slot and the 2^3-block word are both kept per brick cell, so
        // the two loads go out together (cells are below 2^30)
        const uint32_t cell =
            KIND == 3 ? c.node * 4096u + lin3u((x >> 3) & 15u, (y >> 3) & 15u, (z >> 3) & 15u, 16u)
                      : lin3u(x >> 3, y >> 3, z >> 3, (uint32_t)n >> 3);
        c.slot = s.slots[cell];
        c.sub = s.blocks[cell];
        if (COUNT) cnt.n_brickmask++;
      }
    }
    // inside a non-empty brick: step over its empty 4^3 and 2^3 blocks
    // (computed on every step without a branch: outside a brick c.sub is
    // stale and e4, e2 are not used)
    const bool in_brick = !coarse && c.slot != kEmpty;
    // bit of the cell's 2^3 block in sub ((x&7)>>1) + 4 ((y&7)>>1) + 16
    // ((z&7)>>1); its bits 1, 3 and 5 are the 4^3 block's corner bit
    const uint32_t i2 = ((x >> 1) & 3u) | ((y << 1) & 12u) | ((z << 3) & 48u);
    const bool e4 = ((uint32_t)(c.sub >> (i2 & 0x2Au)) & 0x00330033u) == 0;
    const bool e2 = ((c.sub >> i2) & 1ull) == 0;
    if (in_brick && !e4 && !e2) {
      if (COUNT) cnt.n_vox++;
      // 32-bit word index: a scene holds fewer than 2^29 bricks (checked
      // when it is generated or uploaded), so slot * 8 + 7 is below 2^32
      const uint32_t lz = z & 7u, lxy = (x & 7u) | ((y & 7u) << 3);
      const uint64_t m = s.masks[c.slot * 8u + lz];
      if ((m >> lxy) & 1ull) {
        material = s.voxels[(uint64_t)c.slot * 512u + lxy + 64u * lz];
        return kWalkHit;
      }
    }
    c.vmask = c.e128 ? 127u : (c.e32 ? 31u : 7u);
    const int32_t sh =
        c.e128 ? 7 : (c.e32 ? 5 : (!in_brick ? 3 : (e4 ? 2 : (e2 ? 1 : 0))));
    const bool on = advance(r, s, sh, w) && w.t < tmax;
    c.dv = (x ^ (uint32_t)w.cell[0]) | (y ^ (uint32_t)w.cell[1]) | (z ^ (uint32_t)w.cell[2]);
    return on ? kWalkOn : kWalkOut;
  }
}

// Walk until a solid voxel is entered (true), the ray leaves the scene or the
// entry t reaches tmax (false).
template <int KIND, bool COUNT>
__device__ __forceinline__ bool walk(const DevScene &s, const Ray &r, WalkState &w, float tmax,
                                     uint16_t &material, StepCounts &cnt) {
  if (!(w.t < tmax)) return false;
  WalkCache c;
  c.reset();
  for (;;) {
    const int st = walk_step<KIND, COUNT>(s, r, w, c, tmax, material, cnt);
    if (st != kWalkOn) return st == kWalkHit;
  }
}

__device__ __forceinline__ void primary_ray(const DevCamera &c, int32_t px, int32_t py, float o[3],
                                            float d[3]) {
  float a = (float)px + 0.5f;
  a = a * c.inv_w2;
  a = a - 1.0f;
  a = a * c.sx;
  float b = (float)py + 0.5f;
  b = b * c.inv_h2;
  b = 1.0f - b;
  b = b * c.sy;
  float v[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float ra = c.right[i] * a;
    float ub = c.up[i] * b;
    v[i] = c.fwd[i] + ra;
    v[i] = v[i] + ub;
    o[i] = c.eye[i];
  }
  float xx = v[0] * v[0], yy = v[1] * v[1], zz = v[2] * v[2];
  float s = xx + yy;
  s = s + zz;
  float len = __builtin_sqrtf(s);
#pragma unroll
  for (int i = 0; i < 3; i++) d[i] = v[i] / len;
}

struct Hit {
  int64_t index;
  int32_t face;
  float t;
  int32_t cell[3];
  uint16_t material;
};

// Set up the ray o + t d, clip it to the scene's solid box [lo, hi) and find
// its entry cell; false when it misses the box.
__device__ __forceinline__ bool trace_begin(const DevScene &s, const float o[3], const float d[3],
                                            Ray &r, WalkState &w) {
  ray_setup(r, o, d);
  float t_enter = 0.0f, t_exit = __builtin_inff();
  int ea = -1;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (s.lo[a] >= s.hi[a]) return false;  // nothing solid
    const float bl = s.flo[a], bh = s.fhi[a];
    if (r.step[a] == 0) {
      if (o[a] < bl || o[a] >= bh) return false;
      continue;
    }
    float lo = plane_t(r.o[a], r.inv[a], r.step[a] > 0 ? bl : bh);
    float hi = plane_t(r.o[a], r.inv[a], r.step[a] > 0 ? bh : bl);
    if (lo > t_enter) {
      t_enter = lo;
      ea = a;
    }
    if (hi < t_exit) t_exit = hi;
  }
  if (t_enter > t_exit) return false;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (a == ea) {
      w.cell[a] = r.step[a] > 0 ? s.lo[a] : s.hi[a] - 1;
    } else {
      float dt = d[a] * t_enter;
      float p = o[a] + dt;
      int32_t c = (int32_t)__builtin_floorf(p);
      c = c < s.lo[a] ? s.lo[a] : c;
      c = c > s.hi[a] - 1 ? s.hi[a] - 1 : c;
      w.cell[a] = c;
    }
  }
  w.t = t_enter;
  int est = 0;
#pragma unroll
  for (int a = 0; a < 3; a++)
    if (a == ea) est = r.step[a];
  w.face = ea >= 0 ? entry_face(ea, est) : kFaceInside;
  return true;
}

// Clip to [0,n]^3, find the entry cell and walk.
template <int KIND, bool COUNT>
__device__ __forceinline__ bool trace(const DevScene &s, const float o[3], const float d[3],
                                      Hit &hit, StepCounts &cnt) {
  hit.index = -1;
  hit.face = kFaceMiss;
  hit.t = __builtin_inff();
  hit.material = 0;
  Ray r;
  WalkState w;
  if (!trace_begin(s, o, d, r, w)) return false;
  const int32_t n = s.n;
  uint16_t m = 0;
  if (!walk<KIND, COUNT>(s, r, w, __builtin_inff(), m, cnt)) return false;
  hit.index = lin3(w.cell[0], w.cell[1], w.cell[2], n);
  hit.face = w.face;
  hit.t = w.t;
  hit.cell[0] = w.cell[0];
  hit.cell[1] = w.cell[1];
  hit.cell[2] = w.cell[2];
  hit.material = m;
  return true;
}

// true when a secondary ray can start in cell `start` (inside the scene's
// solid box).  A start outside it is the neighbour across a face of a voxel
// on the box's side, and the ray travels into that face's half-space, away
// from the box, so it is unoccluded.
__device__ __forceinline__ bool start_inside(const DevScene &s, const int32_t start[3]) {
  return start[0] >= s.lo[0] && start[0] < s.hi[0] && start[1] >= s.lo[1] && start[1] < s.hi[1] &&
         start[2] >= s.lo[2] && start[2] < s.hi[2];
}

// A secondary ray o + t d that walks from cell `start` at t = 0.
__device__ __forceinline__ void occluded_begin(const float o[3], const float d[3],
                                               const int32_t start[3], Ray &r, WalkState &w) {
  ray_setup(r, o, d);
  w.cell[0] = start[0];
  w.cell[1] = start[1];
  w.cell[2] = start[2];
  w.t = 0.0f;
  w.face = kFaceInside;
#pragma unroll
  for (int a = 0; a < 3; a++)
    r.up[a] = r.step[a] > 0 || (r.step[a] == 0 && (int32_t)__builtin_floorf(o[a]) > start[a]);
}

template <int KIND, bool COUNT>
__device__ __forceinline__ bool occluded(const DevScene &s, const float o[3], const float d[3],
                                         const int32_t start[3], float tmax, StepCounts &cnt) {
  if (!start_inside(s, start)) return false;
  Ray r;
  WalkState w;
  occluded_begin(o, d, start, r, w);
  uint16_t m;
  return walk<KIND, COUNT>(s, r, w, tmax, m, cnt);
}

// lowbias32 integer finaliser (C. Wellons, hash-prospector)
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// Cosine-weighted hemisphere direction about the face normal (Malley's
// method): a point uniform in the unit disk, found by rejection from
// counter-based hashes, lifted by z = sqrt(1 - r^2).
__device__ __forceinline__ void ao_direction(uint32_t seed, uint32_t pixel, uint32_t sample,
                                             int face, float dir[3]) {
  float a = 0.0f, b = 0.0f, r2 = 0.0f;
  const uint32_t base = hash32(hash32(seed) ^ pixel);
  bool ok = false;
  for (uint32_t k = 0; k < 16u; k++) {
    uint32_t h = hash32(base ^ (sample * 16u + k));
    a = (float)(h & 0xFFFFu) * (1.0f / 32768.0f);
    a = a - 1.0f;
    b = (float)(h >> 16) * (1.0f / 32768.0f);
    b = b - 1.0f;
    float aa = a * a, bb = b * b;
    r2 = aa + bb;
    if (r2 < 1.0f) {
      ok = true;
      break;
    }
  }
  if (!ok) {
    a = 0.0f;
    b = 0.0f;
    r2 = 0.0f;
  }
  float z = __builtin_sqrtf(1.0f - r2);
  const int na = face_axis(face);
  const bool pos = (face & 1) == 0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int ua = na == 0 ? 1 : 0;
    const int va = na == 2 ? 1 : 2;
    if (k == na) dir[k] = pos ? z : -z;
    else if (k == ua) dir[k] = a;
    else if (k == va) dir[k] = b;
  }
}

struct PixelResult {
  int64_t index;
  int32_t face;
  float t;
  float shade;
};

// Where the secondary rays of a primary hit start: the hit point p = o + t d
// on entry face `face` (< 6) of voxel `cell`, the cell in front of that face,
// and the face's Lambert term max(0, n . light).
struct Secondary {
  float p[3];
  int32_t start[3];
  float lambert;
};

__device__ __forceinline__ void secondary_setup(const DevShading &sh, const float o[3],
                                                const float d[3], int32_t face, float t,
                                                const int32_t cell[3], Secondary &sc) {
  const int na = face_axis(face);
  const int sg = (face & 1) ? -1 : 1;
  float lightc = sh.light[0];
  sc.start[0] = cell[0];
  sc.start[1] = cell[1];
  sc.start[2] = cell[2];
#pragma unroll
  for (int k = 0; k < 3; k++)
    if (k == na) {
      lightc = sh.light[k];
      sc.start[k] += sg;
    }
  float lambert = sg > 0 ? lightc : -lightc;
  if (lambert < 0.0f) lambert = 0.0f;
  sc.lambert = lambert;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    float dt = d[i] * t;
    sc.p[i] = o[i] + dt;
  }
}

// (0.25 + 0.75 * lambert * visible) * (0.5 + 0.5 * unoccluded / samples),
// the second factor only with AO rays
__device__ __forceinline__ float shade_value(float lambert, float vis, bool ao_on, int32_t unocc,
                                             int32_t samples) {
  float ao = 1.0f;
  if (ao_on) {
    float frac = (float)unocc / (float)samples;
    frac = 0.5f * frac;
    ao = 0.5f + frac;
  }
  float l = lambert * vis;
  l = 0.75f * l;
  float shade = 0.25f + l;
  shade = shade * ao;
  return shade;
}

// The secondary rays of pixel `pix` whose primary ray hit face `face` (< 6),
// set up by secondary_setup: the shadow ray toward the light when the face is
// lit, the AO rays, and the shade.  With neither flag set it traces nothing.
template <int KIND, bool COUNT>
__device__ __forceinline__ float shade_secondary(const DevScene &s, uint32_t flags,
                                                 const DevShading &sh, uint32_t pix, int32_t face,
                                                 const Secondary &sc, StepCounts &cnt,
                                                 uint32_t &n_shadow, uint32_t &n_ao) {
  float vis = 1.0f;
  if ((flags & MESO_RAY_SHADOW) && sc.lambert > 0.0f) {
    n_shadow++;
    if (occluded<KIND, COUNT>(s, sc.p, sh.light, sc.start, __builtin_inff(), cnt)) vis = 0.0f;
  }
  int32_t unocc = 0;
  if (flags & MESO_RAY_AO) {
    for (int32_t k = 0; k < sh.ao_samples; k++) {
      float dir[3];
      ao_direction(sh.ao_seed, pix, (uint32_t)k, face, dir);
      n_ao++;
      if (!occluded<KIND, COUNT>(s, sc.p, dir, sc.start, sh.ao_radius, cnt)) unocc++;
    }
  }
  return shade_value(sc.lambert, vis, (flags & MESO_RAY_AO) != 0, unocc, sh.ao_samples);
}

// Secondary rays of a primary hit (voxel `cell`, entry face `face` < 6, at t
// along o + t d) and the shade.
template <int KIND, bool COUNT>
__device__ __forceinline__ float shade_hit(const DevScene &s, uint32_t flags, const DevShading &sh,
                                           uint32_t pix, const float o[3], const float d[3],
                                           int32_t face, float t, const int32_t cell[3],
                                           StepCounts &cnt, uint32_t &n_shadow, uint32_t &n_ao) {
  Secondary sc;
  secondary_setup(sh, o, d, face, t, cell, sc);
  return shade_secondary<KIND, COUNT>(s, flags, sh, pix, face, sc, cnt, n_shadow, n_ao);
}

// true when a primary hit on `face` needs secondary rays
__device__ __forceinline__ bool needs_secondary(uint32_t flags, const DevShading &sh, int32_t face) {
  if (face >= 6) return false;
  if (flags & MESO_RAY_AO) return true;
  if (!(flags & MESO_RAY_SHADOW)) return false;
  const int na = face_axis(face);
  const float lightc = na == 0 ? sh.light[0] : (na == 1 ? sh.light[1] : sh.light[2]);
  const float lambert = (face & 1) ? -lightc : lightc;
  return lambert > 0.0f;
}

// Primary ray of pixel (px,py), then its secondary rays and shade.
template <int KIND, bool COUNT>
__device__ __forceinline__ PixelResult shade_pixel(const DevScene &s, const DevCamera &cam,
                                                   uint32_t flags, const DevShading &sh,
                                                   int32_t px, int32_t py, StepCounts &cnt,
                                                   uint32_t &n_shadow, uint32_t &n_ao) {
  float o[3], d[3];
  primary_ray(cam, px, py, o, d);
  Hit h;
  PixelResult res;
  const bool hit = trace<KIND, COUNT>(s, o, d, h, cnt);
  float shade = 0.0f;
  if (hit) {
    if (h.face < 6)
      shade = shade_hit<KIND, COUNT>(s, flags, sh, (uint32_t)((int64_t)py * cam.width + px), o, d,
                                     h.face, h.t, h.cell, cnt, n_shadow, n_ao);
    else
      shade = 0.25f;
  }
  res.index = h.index;
  res.face = h.face;
  res.t = h.t;
  res.shade = shade;
  return res;
}

}  // namespace meso
