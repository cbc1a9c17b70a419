// Culled-face and greedy meshing of 32^3 chunks on gfx950.
//
// One workgroup (256 lanes) meshes one chunk:
//  1. Row masks.  Each lane loads 8 voxels of a 32-voxel row with one vector
//     load, plus the same 8 voxels of the +y row and of the +z row, and forms
//     8-bit masks in registers: occupancy and material equality with the +x,
//     +y and +z neighbours.  Four lanes hold a row; two xor-shuffles OR their
//     pieces into 32-bit row masks O[z][y], Ex, Ey, Ez (bits over x) in LDS.
//     A ballot transpose gives the masks with bits over y (OT, EyT, EzT) for
//     the X-facing planes.
//  2. One lane per (face, slice) plane (6 x 32 = 192 lanes) meshes its 32x32
//     plane with a row sweep in bit arithmetic.  F[v] are the exposed faces of
//     row v, EH[v] / EV[v] the material equality along u / v.  The quads that
//     reach row v form disjoint segments (mask `open`, start bits `starts`);
//     a segment continues into row v+1 iff it lies inside
//     G = F[v+1] & EV[v] (all its faces exist there with the same material),
//     tested for all segments at once with a segmented prefix-OR of the
//     failing bits.  The cells of row v+1 not covered by continuing quads
//     split into maximal equal-material runs, each one a new quad.  This is
//     the canonical greedy merge (scan row-major, width first, then height):
//     a quad's height never depends on cells claimed by other quads, because
//     quads created before it never share its columns below its first row.
//     Counting is popcounts of the new-run starts; emitting keeps one packed
//     record (first row, canonical index) per open quad and writes each quad
//     when its segment stops.
//  3. Batched mode: chunks come from an atomic ticket, and each chunk finds
//     its output offset by an ordered decoupled look-back over the chunks
//     ticketed before it, so one persistent launch writes one compact quad
//     buffer plus per-chunk counts and exclusive prefixes.
#include <cstddef>
#include <type_traits>

#include "meso_internal.h"

namespace meso {

constexpr int kPlanes = 192;  // 6 faces x 32 slices
constexpr int kMeshBlock = 256;

struct ChunkAddr {
  const void *base;
  uint64_t off;     // element offset of voxel (0,0,0)
  uint64_t sy, sz;  // element strides of y and z
};

__device__ __forceinline__ ChunkAddr chunk_addr(const meso_chunk_source &src, uint32_t c) {
  ChunkAddr a;
  a.base = src.voxels;
  if (src.layout == MESO_CHUNKS_DENSE_GRID) {
    const uint32_t nc = (uint32_t)src.grid_n / 32u;
    const uint64_t N = (uint64_t)src.grid_n;
    const uint32_t cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
    a.off = (uint64_t)cx * 32u + (uint64_t)cy * 32u * N + (uint64_t)cz * 32u * N * N;
    a.sy = N;
    a.sz = N * N;
  } else {
    a.off = (uint64_t)c * 32768u;
    a.sy = 32;
    a.sz = 1024;
  }
  return a;
}

// 8 consecutive voxels along x as two (u8) or four (u16) packed words.
template <int ES>
struct Vox8 {
  uint32_t w[ES * 2];
  __device__ __forceinline__ uint32_t get(int j) const {
    if constexpr (ES == 1)
      return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu;
    else
      return (w[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
  }
};

template <int ES>
__device__ __forceinline__ Vox8<ES> load8(const ChunkAddr &a, int x0, int y, int z) {
  Vox8<ES> r;
  const uint64_t i = a.off + (uint64_t)x0 + (uint64_t)y * a.sy + (uint64_t)z * a.sz;
  if constexpr (ES == 1) {
    const uint2 v = *(const uint2 *)((const uint8_t *)a.base + i);
    r.w[0] = v.x;
    r.w[1] = v.y;
  } else {
    const uint4 v = *(const uint4 *)((const uint16_t *)a.base + i);
    r.w[0] = v.x;
    r.w[1] = v.y;
    r.w[2] = v.z;
    r.w[3] = v.w;
  }
  return r;
}

// Per 16-bit half of a: min(half, 1), i.e. 1 where the half is nonzero (one
// v_pk_min_u16; written as a generic vector min, the compiler turns it into a
// compare and select per half).
__device__ __forceinline__ uint32_t halves_nonzero(uint32_t a) {
  uint32_t r;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(r) : "v"(a), "v"(0x00010001u));
  return r;
}

// Bit j set where voxel j of four u16 words (voxel 2i in the low half of
// a[i], 2i + 1 in the high half) is nonzero.  halves_nonzero leaves a 1 in
// bit 0 and bit 16 of each word; the four words are shifted together so the
// low halves land on bits 0, 2, 4, 6 and the high halves on bits 16, 18, 20,
// 22, and a shift by 15 takes the latter to bits 1, 3, 5, 7.
__device__ __forceinline__ uint32_t nonzero8(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3) {
  const uint32_t g = halves_nonzero(a0) | (halves_nonzero(a1) << 2) |
                     (halves_nonzero(a2) << 4) | (halves_nonzero(a3) << 6);
  return (g | (g >> 15)) & 0xFFu;
}

// The four mask pieces of 8 u16 voxels (words w) at x = 8k .. 8k + 7:
// occupancy, and equality with the next voxel along x (nx: the first word of
// x + 8), y (wy) and z (wz).  The x pairs (v[j], v[j + 1]) are word i against
// the funnel shift of (w[i + 1]:w[i]) by 16.
struct Pieces {
  uint32_t o, ex, ey, ez;
};

__device__ __forceinline__ Pieces pieces_u16(const uint32_t w[4], uint32_t nx,
                                                const uint32_t wy[4], const uint32_t wz[4],
                                                bool last_k, bool y_in, bool z_in) {
  const uint32_t s0 = (w[0] >> 16) | (w[1] << 16), s1 = (w[1] >> 16) | (w[2] << 16),
                 s2 = (w[2] >> 16) | (w[3] << 16), s3 = (w[3] >> 16) | (nx << 16);
  Pieces p;
  p.o = nonzero8(w[0], w[1], w[2], w[3]);
  p.ex = ~nonzero8(w[0] ^ s0, w[1] ^ s1, w[2] ^ s2, w[3] ^ s3) & (last_k ? 0x7Fu : 0xFFu);
  p.ey = ~nonzero8(w[0] ^ wy[0], w[1] ^ wy[1], w[2] ^ wy[2], w[3] ^ wy[3]) & (y_in ? 0xFFu : 0u);
  p.ez = ~nonzero8(w[0] ^ wz[0], w[1] ^ wz[1], w[2] ^ wz[2], w[3] ^ wz[3]) & (z_in ? 0xFFu : 0u);
  return p;
}

template <int ES>
__device__ __forceinline__ Vox8<ES> zero8() {
  Vox8<ES> r;
#pragma unroll
  for (int k = 0; k < ES * 2; k++) r.w[k] = 0;
  return r;
}

// Rows of a 32^3 chunk of a resident brick scene (KIND 2 brick map, 3
// two-level): the 8 voxels x0..x0+7 of a row are one 16-byte row of a brick.
template <int KIND>
struct SceneRows {
  DevScene s;
  int32_t ox, oy, oz;  // voxel origin of the chunk
};

// Slot of brick (bx, by, bz) of a scene, 0xFFFFFFFF when the brick is empty
// or outside the scene.
template <int KIND>
__device__ __forceinline__ uint32_t scene_slot(const DevScene &s, int32_t bx, int32_t by,
                                               int32_t bz) {
  const int32_t nb = s.n >> 3;
  if (bx >= nb || by >= nb || bz >= nb) return 0xFFFFFFFFu;
  if constexpr (KIND == 2) {
    return s.slots[(uint32_t)bx + ((uint32_t)by + (uint32_t)bz * (uint32_t)nb) * (uint32_t)nb];
  } else {
    const uint32_t nt = (uint32_t)s.n >> 7;
    const uint32_t node =
        s.top[((uint32_t)bx >> 4) + (((uint32_t)by >> 4) + ((uint32_t)bz >> 4) * nt) * nt];
    if (node == 0xFFFFFFFFu) return 0xFFFFFFFFu;
    return s.slots[(uint64_t)node * 4096u + ((uint32_t)bx & 15u) + 16u * ((uint32_t)by & 15u) +
                   256u * ((uint32_t)bz & 15u)];
  }
}

// Rows of a scene chunk whose 4x4x4 brick slots are resolved: bslot[lx + 4(ly
// + 4 lz)] for the chunk's local brick (lx, ly, lz).  Every row the mask
// build reads lies in the chunk, so a row is one 16-byte load.
struct SlotRows {
  const uint32_t *bslot;
  const uint16_t *voxels;
};

template <int ES>
__device__ __forceinline__ Vox8<2> load8(const SlotRows &a, int x0, int y, int z) {
  Vox8<2> r = zero8<2>();
  const uint32_t slot = a.bslot[(x0 >> 3) + 4 * (y >> 3) + 16 * (z >> 3)];
  if (slot == 0xFFFFFFFFu) return r;
  const uint4 v = *(const uint4 *)(a.voxels + (uint64_t)slot * 512u + 8u * (uint32_t)(y & 7) +
                                   64u * (uint32_t)(z & 7));
  r.w[0] = v.x;
  r.w[1] = v.y;
  r.w[2] = v.z;
  r.w[3] = v.w;
  return r;
}

constexpr int kZY = 33;
enum : int {
  kO = 0,
  kEx = 32 * kZY,
  kEy = 64 * kZY,
  kEz = 96 * kZY,
  kOT = 128 * kZY,
  kEyT = kOT + 1024,
  kEzT = kOT + 2048,
  kMaskWords = kOT + 3072
};

struct MeshLds {
  uint32_t M[kMaskWords];
  uint32_t pex[kPlanes];  // culled: exclusive quad offset of each plane in the chunk
  uint32_t occ[kPlanes];  // culled: bit v set when row v of the plane has a face
  uint32_t scan[kMeshBlock / 64];
  uint32_t bslot[64];  // scene chunks: slots of the chunk's 4x4x4 bricks
  uint32_t chunk;
  uint32_t base_lo, base_hi;
};

// 32x32 bit transpose across the 32 lanes of a half wave: lane y holds row y
// (bit x = element [y][x]) and gets back column y (bit x = element [x][y]).
// Stage j swaps the high-j bits of lane y (y & j == 0) with the low-j bits of
// lane y + j.  Both halves' results are formed and one is kept by a bit
// select: written as a conditional, the compiler branched on the lane's bit
// (both sides ran anyway, plus the exec-mask handling).
__device__ __forceinline__ uint32_t transpose32(uint32_t w, uint32_t y) {
  constexpr uint32_t hi[5] = {0xFFFF0000u, 0xFF00FF00u, 0xF0F0F0F0u, 0xCCCCCCCCu, 0xAAAAAAAAu};
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const int j = 16 >> k;
    const uint32_t p = (uint32_t)__shfl_xor((int)w, j, 32);
    const uint32_t h = hi[k];
    const uint32_t up = (w & h) | ((p >> j) & ~h);
    const uint32_t dn = (w & ~h) | ((p << j) & h);
    const uint32_t m = 0u - ((y >> (4 - k)) & 1u);  // all ones where y & j
    w = (up & m) | (dn & ~m);
  }
  return w;
}

// OR of the 8-bit pieces held by the 4 lanes of a row (lanes 4r .. 4r+3)
__device__ __forceinline__ uint32_t row_or(uint32_t piece) {
  piece |= __shfl_xor(piece, 1, 64);
  piece |= __shfl_xor(piece, 2, 64);
  return piece;
}

// Step 1: masks.  Wave w handles blocks of 16 rows (rows r = y + 32 z).
template <int ES, class Rows>
__device__ __forceinline__ void build_masks(MeshLds &L, const Rows &a) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  constexpr uint32_t nwaves = kMeshBlock / 64;
  const int k = (int)(lane & 3u);
  // two row blocks per trip: all their loads are issued before either
  // block's arithmetic, so a trip waits on memory once
  // (scene chunks only: the batched packed-chunk meshers hide latency with
  // occupancy, and measured slower with the extra registers)
  constexpr uint32_t kTrips = 64u / nwaves, kPer = std::is_same<Rows, SlotRows>::value ? 2u : 1u;
  for (uint32_t it = 0; it < kTrips; it += kPer) {
    Vox8<ES> c[kPer], cy[kPer], cz[kPer];
#pragma unroll
    for (uint32_t b = 0; b < kPer; b++) {
      const uint32_t r = (wave + (it + b) * nwaves) * 16u + (lane >> 2);
      const int y = (int)(r & 31u), z = (int)(r >> 5);
      c[b] = load8<ES>(a, 8 * k, y, z);
      // the row itself stands in past the chunk's edge (masked off below)
      cy[b] = load8<ES>(a, 8 * k, y < 31 ? y + 1 : y, z);
      cz[b] = load8<ES>(a, 8 * k, y, z < 31 ? z + 1 : z);
    }
#pragma unroll
    for (uint32_t b = 0; b < kPer; b++) {
      const uint32_t r = (wave + (it + b) * nwaves) * 16u + (lane >> 2);
      const int y = (int)(r & 31u), z = (int)(r >> 5);
      const uint32_t next0 = __shfl_down(c[b].w[0], 1, 64);  // first voxels of x + 8
      uint32_t o = 0, ex = 0, ey = 0, ez = 0;
      if constexpr (ES == 2) {
        const Pieces p = pieces_u16(c[b].w, next0, cy[b].w, cz[b].w, k == 3, y < 31, z < 31);
        o = p.o;
        ex = p.ex;
        ey = p.ey;
        ez = p.ez;
      } else {
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const uint32_t v = c[b].get(j);
          o |= (uint32_t)(v != 0u) << j;
          const uint32_t vx = j < 7 ? c[b].get(j + 1) : next0 & 0xFFu;
          ex |= (uint32_t)(v == vx && (j < 7 || k < 3)) << j;
          ey |= (uint32_t)(y < 31 && v == cy[b].get(j)) << j;
          ez |= (uint32_t)(z < 31 && v == cz[b].get(j)) << j;
        }
      }
      const int sh = 8 * k;
      o = row_or(o << sh);
      ex = row_or(ex << sh);
      ey = row_or(ey << sh);
      ez = row_or(ez << sh);
      if (k == 0) {
        const int zy = z * kZY + y;
        L.M[kO + zy] = o;
        L.M[kEx + zy] = ex;
        L.M[kEy + zy] = ey;
        L.M[kEz + zy] = ez;
      }
    }
  }
  __syncthreads();
  // transposes: lane y + 32 * (z & 1) holds row [z][y] (bits over x) of two z
  // planes; five block-swap stages between lanes y and y ^ j (j = 16 .. 1)
  // turn it into column [z][x] (bits over y), one per lane
  for (uint32_t zp = wave; zp < 16u; zp += nwaves) {
    const uint32_t z = zp * 2u + (lane >> 5);
    const uint32_t y = lane & 31u;
    const uint32_t rows[3] = {L.M[kO + z * kZY + y], L.M[kEy + z * kZY + y],
                              L.M[kEz + z * kZY + y]};
    constexpr int dst[3] = {kOT, kEyT, kEzT};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      L.M[dst[k] + z * 32u + y] = transpose32(rows[k], y);
    }
  }
  __syncthreads();
}

// Masks of scene chunk (ox, oy, oz) / 32: the slots of its 64 bricks are
// resolved first, one lane each, so that the row loads do not wait on the
// brick lookups.
template <int ES, int KIND>
__device__ __forceinline__ void build_masks(MeshLds &L, const SceneRows<KIND> &a) {
  const uint32_t t = threadIdx.x;
  if (t < 64u)
    L.bslot[t] = scene_slot<KIND>(a.s, (a.ox >> 3) + (int32_t)(t & 3u),
                                  (a.oy >> 3) + (int32_t)((t >> 2) & 3u),
                                  (a.oz >> 3) + (int32_t)(t >> 4));
  __syncthreads();
  build_masks<ES>(L, SlotRows{L.bslot, a.s.voxels});
}

// Offsets of plane (f, s) in the mask planes.  z faces: (u,v) = (x,y) and
// rows are O[s][v]; y faces: (x,z), rows O[v][s]; x faces: (y,z), rows OT[v][s].
struct PlaneView {
  int self, nb, eu, ev, stride;  // nb < 0: the neighbouring slice is outside
};

__device__ __forceinline__ PlaneView plane_view(int f, int s) {
  PlaneView p;
  const int ns = (f & 1) ? s - 1 : s + 1;
  const bool inside = ns >= 0 && ns < 32;
  if (f < 2) {
    p.self = kO + s * kZY;
    p.nb = inside ? kO + ns * kZY : -1;
    p.eu = kEx + s * kZY;
    p.ev = kEy + s * kZY;
    p.stride = 1;
  } else if (f < 4) {
    p.self = kO + s;
    p.nb = inside ? kO + ns : -1;
    p.eu = kEx + s;
    p.ev = kEz + s;
    p.stride = kZY;
  } else {
    p.self = kOT + s;
    p.nb = inside ? kOT + ns : -1;
    p.eu = kEyT + s;
    p.ev = kEzT + s;
    p.stride = 32;
  }
  return p;
}

__device__ __forceinline__ uint32_t face_row(const MeshLds &L, const PlaneView &p, int v) {
  return L.M[p.self + v * p.stride] & ~(p.nb >= 0 ? L.M[p.nb + v * p.stride] : 0u);
}

// face_row at word offset `off` (= v * stride) of the plane's rows, without
// a branch: a plane on the chunk's side reads its own row as the neighbour
// and masks it off.
__device__ __forceinline__ uint32_t face_row_at(const MeshLds &L, const PlaneView &p, int off) {
  const uint32_t nbm = p.nb >= 0 ? ~0u : 0u;
  const int nb = p.nb >= 0 ? p.nb : p.self;
  return L.M[p.self + off] & ~(L.M[nb + off] & nbm);
}

__device__ __forceinline__ uint32_t pack_quad(int f, int s, int u, int v, int w, int h) {
  int x, y, z;
  if (f < 2) {
    x = u; y = v; z = s;
  } else if (f < 4) {
    x = u; y = s; z = v;
  } else {
    x = s; y = u; z = v;
  }
  return (uint32_t)x | ((uint32_t)y << 5) | ((uint32_t)z << 10) | ((uint32_t)(w - 1) << 15) |
         ((uint32_t)(h - 1) << 20) | ((uint32_t)f << 25);
}

// x and every bit reachable from it by moving up through bits of g (bit j+1
// reachable from bit j when bit j+1 is in g).  A seed just above each bit of
// x that lies in g is added to g: the addition turns the run of g's ones from
// the seed to the run's top into zeros and leaves the carry on the first bit
// above the run, which is outside g; so (g + s) ^ g marks the run from the
// seed up (plus that bit, masked off by & g).  A second seed in the same run
// is already inside the flipped stretch except for its own bit, added by | s.
__device__ __forceinline__ uint32_t fill_up(uint32_t x, uint32_t g) {
  const uint32_t s = (x << 1) & g;
  return x | ((((g + s) ^ g) | s) & g);
}

// Bits of the segments that contain a marked bit.  Segments are runs of
// `cont` bits (a bit continuing the segment of the bit below it) above a start
// bit; the marks spread up through `cont`, and down (bit j reached from bit
// j+1 when bit j+1 is in `cont`) by the same fill on the bit-reversed words.
__device__ __forceinline__ uint32_t segment_fill(uint32_t marks, uint32_t cont) {
  const uint32_t up = fill_up(marks, cont);
  const uint32_t dn = __brev(fill_up(__brev(marks), __brev(cont) << 1));
  return up | dn;
}

// Step 2a: quad count of plane t; culled meshing also returns the plane's
// non-empty rows in `occ`.
template <bool GREEDY>
__device__ __forceinline__ uint32_t plane_count(const MeshLds &L, int t, uint32_t &occ) {
  const PlaneView p = plane_view(t >> 5, t & 31);
  uint32_t count = 0;
  occ = 0;
  if (!GREEDY) {
#pragma unroll 8
    for (int v = 0; v < 32; v++) {
      const uint32_t c = (uint32_t)__popc(face_row(L, p, v));
      count += c;
      occ |= (c != 0u ? 1u : 0u) << v;
    }
    return count;
  }
  uint32_t open = 0, starts = 0, ev_prev = 0;
#pragma unroll 4
  for (int v = 0; v < 32; v++) {
    const uint32_t F = face_row_at(L, p, v * p.stride);
    const uint32_t eh = L.M[p.eu + v * p.stride];
    if (v > 0) {
      const uint32_t dead = segment_fill(open & ~(F & ev_prev), open & ~starts) & open;
      open &= ~dead;
      starts &= ~dead;
    }
    const uint32_t avail = F & ~open;
    const uint32_t st = avail & ~((avail & eh) << 1);
    count += (uint32_t)__popc(st);
    open |= avail;
    starts |= st;
    ev_prev = L.M[p.ev + v * p.stride];
  }
  return count;
}

// Start rows of the open quads, bit-sliced: bit u of row[k] is bit k of the
// first row of the quad that starts at u (kept in registers, no table).
struct StartRows {
  uint32_t row[5];
};

__device__ __forceinline__ void record_starts(StartRows &r, uint32_t st, int v) {
#pragma unroll
  for (int k = 0; k < 5; k++) r.row[k] = (r.row[k] & ~st) | (((v >> k) & 1) ? st : 0u);
}

__device__ __forceinline__ uint32_t start_row(const StartRows &r, uint32_t u) {
  uint32_t v = __builtin_amdgcn_ubfe(r.row[0], u, 1u);
#pragma unroll
  for (int k = 1; k < 5; k++) v |= __builtin_amdgcn_ubfe(r.row[k], u, 1u) << k;
  return v;
}

// pack_quad with plane (f, s)'s axes resolved once:
// quad = fixed | u << su | v << sv | (w - 1) << 15 | (h - 1) << 20.
struct PlanePack {
  uint32_t fixed, su, sv;
};

__device__ __forceinline__ PlanePack plane_pack(int f, int s) {
  PlanePack p;
  if (f < 2) {  // (x, y, z) = (u, v, s)
    p.su = 0;
    p.sv = 5;
    p.fixed = (uint32_t)s << 10;
  } else if (f < 4) {  // (u, s, v)
    p.su = 0;
    p.sv = 10;
    p.fixed = (uint32_t)s << 5;
  } else {  // (s, u, v)
    p.su = 5;
    p.sv = 10;
    p.fixed = (uint32_t)s;
  }
  p.fixed |= (uint32_t)f << 25;
  return p;
}

// The quad of the segment that starts at u (its columns: u and the `cont`
// bits above it), first row v, last row vend - 1.  (cont >> u) >> 1 is 0 at
// u = 31, where the width is 1.
__device__ __forceinline__ uint32_t closed_quad(const PlanePack &pk, uint32_t cont, uint32_t u,
                                                uint32_t v, uint32_t vend) {
  const uint32_t wm1 = (uint32_t)__builtin_ctz(~((cont >> u) >> 1));
  return pk.fixed | (u << pk.su) | (v << pk.sv) | (wm1 << 15) | ((vend - v - 1u) << 20);
}

// Close the quads whose start bits are in `bits` (their last row is
// vend - 1), in ascending u, at out[base + count ...]; CHECK: stop writing at
// cap (a caller that knows the plane fits passes false and skips the tests).
template <bool CHECK>
__device__ __forceinline__ void close_quads(const StartRows &r, const PlanePack &pk, uint32_t bits,
                                            uint32_t open, uint32_t starts, uint32_t vend,
                                            uint32_t *out, uint64_t base, uint64_t cap,
                                            uint32_t &count) {
  const uint32_t cont = open & ~starts;
  while (bits) {  // two quads per trip
    const uint32_t u1 = (uint32_t)__builtin_ctz(bits);
    bits &= bits - 1;
    const bool two = bits != 0u;
    const uint32_t u2 = two ? (uint32_t)__builtin_ctz(bits) : u1;
    bits &= two ? bits - 1 : bits;
    const uint32_t q1 = closed_quad(pk, cont, u1, start_row(r, u1), vend);
    const uint32_t q2 = closed_quad(pk, cont, u2, start_row(r, u2), vend);
    const uint64_t k = base + count;
    count += two ? 2u : 1u;
    if (!CHECK || k < cap) out[k] = q1;
    if (two && (!CHECK || k + 1 < cap)) out[k + 1] = q2;
  }
}

// Step 2b, greedy: write the quads of plane t at out[base ...] in canonical
// order: by the row where they end, then by start u (the order the sweep
// closes them in).
template <bool CHECK>
__device__ __forceinline__ void plane_emit_greedy(MeshLds &L, int t, uint32_t *out, uint64_t base,
                                                  uint64_t cap) {
  const int f = t >> 5, s = t & 31;
  const PlaneView p = plane_view(f, s);
  const PlanePack pk = plane_pack(f, s);
  uint32_t count = 0;
  uint32_t open = 0, starts = 0, ev_prev = 0;
  StartRows r = {{0u, 0u, 0u, 0u, 0u}};
  int off = 0;  // v * stride, kept as a running sum (no multiply per row)
  for (int v = 0; v < 32; v++, off += p.stride) {
    const uint32_t F = face_row_at(L, p, off);
    const uint32_t eh = L.M[p.eu + off];
    const uint32_t ev = L.M[p.ev + off];
    if (v > 0) {
      const uint32_t dead = segment_fill(open & ~(F & ev_prev), open & ~starts) & open;
      close_quads<CHECK>(r, pk, starts & dead, open, starts, (uint32_t)v, out, base, cap, count);
      open &= ~dead;
      starts &= ~dead;
    }
    const uint32_t avail = F & ~open;
    const uint32_t st = avail & ~((avail & eh) << 1);
    open |= avail;
    starts |= st;
    record_starts(r, st, v);
    ev_prev = ev;
  }
  close_quads<CHECK>(r, pk, starts, open, starts, 32u, out, base, cap, count);
}

// Step 2b, greedy: the quads of plane t (`total` of them, from plane_count)
// at out[base ...], written unchecked when they all fit below cap.  (Culled
// meshing shares its faces out over the whole workgroup instead:
// culled_emit_balanced.)
__device__ __forceinline__ void plane_emit(MeshLds &L, int t, uint32_t *out, uint64_t base,
                                           uint64_t cap, uint32_t total) {
  if (base + total <= cap)
    plane_emit_greedy<false>(L, t, out, base, cap);
  else
    plane_emit_greedy<true>(L, t, out, base, cap);
}

// position of the r-th (0-based) set bit of x; x has more than r set bits
__device__ __forceinline__ int select_bit(uint32_t x, uint32_t r) {
  int pos = 0;
#pragma unroll
  for (int w = 16; w >= 1; w >>= 1) {
    const uint32_t c = (uint32_t)__popc(x & ((1u << w) - 1u));
    if (r >= c) {
      r -= c;
      x >>= w;
      pos += w;
    }
  }
  return pos;
}

// the last plane whose quad offset is <= q (offsets ascend and repeat over
// empty planes, so for q < total this is the plane holding quad q)
__device__ __forceinline__ int plane_of(const MeshLds &L, uint32_t q) {
  int lo = 0, hi = kPlanes - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (L.pex[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Culled emission spread over the whole workgroup in runs of kRun quads:
// thread i writes the runs [kRun (i + n k), kRun (i + n k + 1)) of the chunk
// (n = workgroup size, k = 0, 1, ...), each in canonical order (plane, row,
// u), so one store instruction covers n * kRun consecutive quads (coalesced)
// and every thread loops about the same number of times whatever the spread
// of faces over the planes.  A thread finds the first face of a run by a
// binary search over the planes' offsets (pex) and a walk over the plane's
// non-empty rows (occ); it crosses to the next plane by the same search.
constexpr uint32_t kRun = 8;

__device__ __forceinline__ void culled_emit_balanced(const MeshLds &L, uint32_t total,
                                                     uint32_t *out, uint64_t base, uint64_t cap) {
  for (uint32_t q0 = threadIdx.x * kRun; q0 < total; q0 += kMeshBlock * kRun) {
    const uint32_t q1 = q0 + kRun < total ? q0 + kRun : total;
    int pl = plane_of(L, q0);
    PlaneView p = plane_view(pl >> 5, pl & 31);
    uint32_t r = q0 - L.pex[pl];
    uint32_t rest = L.occ[pl];  // non-empty rows not yet entered
    int v = __ffs(rest) - 1;
    rest &= rest - 1u;
    uint32_t row = face_row(L, p, v);
    for (;;) {  // the row holding face r of the plane
      const uint32_t c = (uint32_t)__popc(row);
      if (r < c) break;
      r -= c;
      v = __ffs(rest) - 1;
      rest &= rest - 1u;
      row = face_row(L, p, v);
    }
    row &= ~((1u << select_bit(row, r)) - 1u);  // drop the faces before q0
    for (uint32_t q = q0;;) {
      const int u = __ffs(row) - 1;
      row &= row - 1u;
      const uint64_t k = base + q;
      if (k < cap) out[k] = pack_quad(pl >> 5, pl & 31, u, v, 1, 1);
      if (++q >= q1) break;
      if (row == 0u) {  // next non-empty row, else the plane holding quad q
        if (rest == 0u) {
          pl = plane_of(L, q);
          p = plane_view(pl >> 5, pl & 31);
          rest = L.occ[pl];
        }
        v = __ffs(rest) - 1;
        rest &= rest - 1u;
        row = face_row(L, p, v);
      }
    }
  }
}

// exclusive scan of one value per lane over the workgroup; returns the total
__device__ __forceinline__ uint32_t block_exclusive_scan(MeshLds &L, uint32_t v, uint32_t &excl) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += o;
  }
  if (lane == 63) L.scan[wave] = inc;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < kMeshBlock / 64; w++) {
    const uint32_t s = L.scan[w];
    if (w < wave) before += s;
    total += s;
  }
  excl = before + inc - v;
  __syncthreads();
  return total;
}

// look-back status word: flag in bits 62-63 (1 = aggregate, 2 = inclusive)
constexpr uint64_t kAgg = 1ull << 62, kInc = 2ull << 62, kValMask = (1ull << 62) - 1;

struct MeshWork {
  unsigned int ticket;
  unsigned int error;
  unsigned int pad[62];
  unsigned long long status[1];  // n_chunks entries
};

// chunks of a resident scene: chunk coordinates on the device
struct SceneSource {
  DevScene s;
  const int32_t *xyz;
  int32_t n_chunks;
};

// SRC 0: packed or dense-grid chunks (meso_chunk_source); 2 / 3: chunks of a
// brick map / two-level scene
template <int ES, bool GREEDY, int SRC>
__global__ void __launch_bounds__(kMeshBlock) mesh_batched_kernel(
    meso_chunk_source src, SceneSource ss, uint32_t *counts, uint32_t *prefix, uint32_t *quads,
    uint64_t capacity, unsigned long long *total, MeshWork *work, uint32_t spin_limit) {
  __shared__ MeshLds L;
  const int t = (int)threadIdx.x;
  const uint32_t n_chunks = (uint32_t)(SRC == 0 ? src.n_chunks : ss.n_chunks);
  for (;;) {
    if (t == 0) L.chunk = atomicAdd(&work->ticket, 1u);
    __syncthreads();
    const uint32_t c = L.chunk;
    if (c >= n_chunks) break;  // uniform: every wave leaves
    if constexpr (SRC == 0) {
      build_masks<ES>(L, chunk_addr(src, c));
    } else {
      const int32_t nc = (ss.s.n + 31) >> 5;  // a partial last chunk counts
      int32_t cx = ss.xyz[3 * c], cy = ss.xyz[3 * c + 1], cz = ss.xyz[3 * c + 2];
      if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc)
        cx = cy = cz = nc;  // outside the scene: every brick lookup reads empty
      build_masks<ES>(L, SceneRows<SRC>{ss.s, cx * 32, cy * 32, cz * 32});
    }
    uint32_t occ = 0;
    const uint32_t mine = t < kPlanes ? plane_count<GREEDY>(L, t, occ) : 0u;
    uint32_t excl = 0;
    const uint32_t chunk_total = block_exclusive_scan(L, mine, excl);
    if (t == 0) {
      // publish the aggregate, then add up predecessors until an inclusive one.
      // The value travels inside the status word and no other data is
      // exchanged, so the status accesses are relaxed: agent-scope release /
      // acquire would write back and invalidate this XCD's whole L2 at every
      // chunk (and every spin), evicting the chunk data of the other
      // workgroups on it.
      __hip_atomic_store(&work->status[c], kAgg | (uint64_t)chunk_total, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      uint64_t before = 0;
      int64_t j = (int64_t)c - 1;
      uint32_t spins = 0;
      while (j >= 0) {
        // chunk j was ticketed earlier, so its workgroup is resident and will
        // publish; the bound (the context's spin limit, 2^24 polls unless the
        // host lowers it to inject the failure) only guarantees that every
        // wave exits.  The chunk's outputs are then wrong and the error word
        // says so (meso_mesh_check).
        if (spins >= spin_limit) {
          atomicOr(&work->error, 1u);
          break;
        }
        const uint64_t st = __hip_atomic_load(&work->status[j], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        if (st & kInc) {
          before += st & kValMask;
          break;
        }
        if (st & kAgg) {
          before += st & kValMask;
          j--;
          continue;
        }
        spins++;
        __builtin_amdgcn_s_sleep(1);
      }
      __hip_atomic_store(&work->status[c], kInc | (before + chunk_total), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
      counts[c] = chunk_total;
      prefix[c] = (uint32_t)before;
      if (c == n_chunks - 1) *total = before + chunk_total;
      L.base_lo = (uint32_t)before;
      L.base_hi = (uint32_t)(before >> 32);
    }
    if constexpr (!GREEDY)
      if (t < kPlanes) {
        L.pex[t] = excl;
        L.occ[t] = occ;
      }
    __syncthreads();
    const uint64_t base = ((uint64_t)L.base_hi << 32) | L.base_lo;
    if constexpr (GREEDY) {
      if (t < kPlanes) plane_emit(L, t, quads, base + excl, capacity, mine);
    } else {
      culled_emit_balanced(L, chunk_total, quads, base, capacity);
    }
    __syncthreads();
  }
}

// Chunks of a resident scene meshed into fixed slots: chunk c of the list
// (coordinates xyz[3c]) writes its quads at slots[c] * per_slot, at most
// per_slot of them, and its full count to slot_count[slots[c]].  The list
// length is read on the device (*n_dev, at most max_n), so that a producer
// kernel can size it.  Workgroups take chunks from ticket[0]; the last workgroup to leave
// (ticket[1] counts leavers) sets both back to zero for the next launch.
template <bool GREEDY, int SRC>
__global__ void __launch_bounds__(kMeshBlock) mesh_slotted_kernel(
    DevScene sc, const int32_t *xyz, const uint32_t *slots, const uint32_t *n_dev, uint32_t max_n,
    uint32_t *slot_count, uint32_t *quads, uint32_t per_slot, unsigned int *ticket) {
  __shared__ MeshLds L;
  const int t = (int)threadIdx.x;
  const uint32_t n_chunks = *n_dev < max_n ? *n_dev : max_n;
  const int32_t nc = (sc.n + 31) >> 5;  // a partial last chunk counts
  for (;;) {
    if (t == 0) L.chunk = atomicAdd(&ticket[0], 1u);
    __syncthreads();
    const uint32_t c = L.chunk;
    if (c >= n_chunks) break;  // uniform: every wave leaves
    int32_t cx = xyz[3 * c], cy = xyz[3 * c + 1], cz = xyz[3 * c + 2];
    if (cx < 0 || cy < 0 || cz < 0 || cx >= nc || cy >= nc || cz >= nc)
      cx = cy = cz = nc;  // outside the scene: every brick lookup reads empty
    build_masks<2>(L, SceneRows<SRC>{sc, cx * 32, cy * 32, cz * 32});
    uint32_t occ = 0;
    const uint32_t mine = t < kPlanes ? plane_count<GREEDY>(L, t, occ) : 0u;
    uint32_t excl = 0;
    const uint32_t chunk_total = block_exclusive_scan(L, mine, excl);
    const uint32_t slot = slots[c];
    const uint64_t base = (uint64_t)slot * per_slot, cap = base + per_slot;
    if (t == 0) slot_count[slot] = chunk_total;
    if constexpr (GREEDY) {
      if (t < kPlanes) plane_emit(L, t, quads, base + excl, cap, mine);
    } else {
      if (t < kPlanes) {
        L.pex[t] = excl;
        L.occ[t] = occ;
      }
      __syncthreads();
      culled_emit_balanced(L, chunk_total, quads, base, cap);
    }
    __syncthreads();
  }
  if (t == 0 && atomicAdd(&ticket[1], 1u) == gridDim.x - 1u) {
    __hip_atomic_store(&ticket[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&ticket[1], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <int ES, bool GREEDY>
__global__ void __launch_bounds__(kMeshBlock) mesh_plain_kernel(const void *voxels, uint32_t *count,
                                                                uint32_t *quads, uint32_t capacity) {
  __shared__ MeshLds L;
  const int t = (int)threadIdx.x;
  const ChunkAddr a{voxels, 0, 32, 1024};
  build_masks<ES>(L, a);
  uint32_t occ = 0;
  const uint32_t mine = t < kPlanes ? plane_count<GREEDY>(L, t, occ) : 0u;
  uint32_t excl = 0;
  const uint32_t chunk_total = block_exclusive_scan(L, mine, excl);
  if (t == 0) *count = chunk_total;
  if constexpr (GREEDY) {
    if (t < kPlanes) plane_emit(L, t, quads, excl, capacity, mine);
  } else {
    if (t < kPlanes) {
      L.pex[t] = excl;
      L.occ[t] = occ;
    }
    __syncthreads();
    culled_emit_balanced(L, chunk_total, quads, 0, capacity);
  }
}

template <int ES, bool GREEDY, int SRC>
static meso_status launch_batched_t(meso_context *ctx, const meso_chunk_source &src,
                                    const SceneSource &ss, int32_t n_chunks, uint32_t *d_counts,
                                    uint32_t *d_prefix, uint32_t *d_quads, uint64_t capacity,
                                    uint64_t *d_total, void *d_work, hipStream_t stream,
                                    const MeshLaunch *shape) {
  int grid = shape && shape->grid > 0 ? shape->grid : 0;
  if (!grid) {
    int per_cu = 0;
    MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, mesh_batched_kernel<ES, GREEDY, SRC>, kMeshBlock, 0));
    if (per_cu < 1) per_cu = 1;
    grid = per_cu * ctx->props.multiProcessorCount;
  }
  if (grid > n_chunks) grid = n_chunks;
  meso_status st = launch_zero(d_work, meso_mesh_workspace_bytes(n_chunks), stream);
  if (st != MESO_OK) return st;
  hipLaunchKernelGGL((mesh_batched_kernel<ES, GREEDY, SRC>), dim3(grid), dim3(kMeshBlock), 0,
                     stream, src, ss, d_counts, d_prefix, d_quads, capacity,
                     (unsigned long long *)d_total, (MeshWork *)d_work,
                     ctx->lookback_spin_limit);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

// Scratch of meso_mesh_chunks_noalloc: the mesh workspace (ticket, error word
// and one look-back status per chunk) for up to MESO_MESH_NOALLOC_MAX_CHUNKS
// chunks, 131,328 bytes, held on each device the library runs on until the
// process ends.  Host code never passes this name to a kernel: it stands for
// a host-side shadow; init_mesh_scratch takes each device's own address.
constexpr uint64_t kMeshScratchBytes = 256u + 8u * (uint64_t)MESO_MESH_NOALLOC_MAX_CHUNKS;
static __device__ __align__(256) unsigned char scratch[kMeshScratchBytes];

meso_status init_mesh_scratch(meso_context *ctx) {
  // called by meso_context_create with ctx->device current
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(MESO_ERR_RUNTIME, "mesher scratch address: %s", hipGetErrorString(e));
  ctx->mesh_scratch = p;
  return MESO_OK;
}

MeshLaunch default_mesh_launch(meso_context *) { return MeshLaunch{kMeshBlock, 0}; }

static meso_status check_source(const meso_chunk_source *src) {
  MESO_CHECK(src && src->size >= sizeof(meso_chunk_source), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_chunk_source missing or too small");
  MESO_CHECK(src->voxels, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "chunk voxels NULL");
  MESO_CHECK(((uintptr_t)src->voxels & 15u) == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "chunk voxels must be 16-byte aligned");
  MESO_CHECK(src->elem_bytes == 1 || src->elem_bytes == 2, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "elem_bytes must be 1 or 2");
  MESO_CHECK(src->n_chunks >= 1 && src->n_chunks <= 40000, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 1..40000 (quad offsets are u32)", src->n_chunks);
  if (src->layout == MESO_CHUNKS_DENSE_GRID) {
    MESO_CHECK(src->grid_n >= 32 && src->grid_n % 32 == 0 && src->grid_n <= 4096,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "grid_n %d must be a multiple of 32", src->grid_n);
    const int64_t nc = src->grid_n / 32;
    MESO_CHECK(src->n_chunks <= nc * nc * nc, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "n_chunks %d exceeds the %lld chunks of the grid", src->n_chunks,
               (long long)(nc * nc * nc));
  } else {
    MESO_CHECK(src->layout == MESO_CHUNKS_PACKED, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
               "unknown chunk layout %d", src->layout);
  }
  return MESO_OK;
}

meso_status launch_mesh_batched(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                uint64_t capacity, uint64_t *d_total, void *d_work,
                                hipStream_t stream, const MeshLaunch *shape) {
  const bool g = mode == MESO_MESH_GREEDY;
  const SceneSource none{};
  const int32_t n = src->n_chunks;
  if (src->elem_bytes == 1)
    return g ? launch_batched_t<1, true, 0>(ctx, *src, none, n, d_counts, d_prefix, d_quads,
                                            capacity, d_total, d_work, stream, shape)
             : launch_batched_t<1, false, 0>(ctx, *src, none, n, d_counts, d_prefix, d_quads,
                                             capacity, d_total, d_work, stream, shape);
  return g ? launch_batched_t<2, true, 0>(ctx, *src, none, n, d_counts, d_prefix, d_quads, capacity,
                                          d_total, d_work, stream, shape)
           : launch_batched_t<2, false, 0>(ctx, *src, none, n, d_counts, d_prefix, d_quads,
                                           capacity, d_total, d_work, stream, shape);
}

static meso_status launch_mesh_scene(meso_context *ctx, const meso_scene *scene,
                                     const int32_t *xyz, int32_t n, uint32_t mode,
                                     uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                     uint64_t capacity, uint64_t *d_total, void *d_work,
                                     hipStream_t stream) {
  const meso_chunk_source none{};
  const SceneSource ss{scene->dev(), xyz, n};
  const bool g = mode == MESO_MESH_GREEDY;
  if (scene->kind == MESO_SCENE_BRICKMAP_U16)
    return g ? launch_batched_t<2, true, 2>(ctx, none, ss, n, d_counts, d_prefix, d_quads,
                                            capacity, d_total, d_work, stream, nullptr)
             : launch_batched_t<2, false, 2>(ctx, none, ss, n, d_counts, d_prefix, d_quads,
                                             capacity, d_total, d_work, stream, nullptr);
  return g ? launch_batched_t<2, true, 3>(ctx, none, ss, n, d_counts, d_prefix, d_quads, capacity,
                                          d_total, d_work, stream, nullptr)
           : launch_batched_t<2, false, 3>(ctx, none, ss, n, d_counts, d_prefix, d_quads,
                                           capacity, d_total, d_work, stream, nullptr);
}

template <bool GREEDY, int SRC>
static meso_status launch_slotted_t(meso_context *ctx, const meso_scene *scene,
                                    const int32_t *xyz, const uint32_t *slots,
                                    const uint32_t *n_dev, int32_t max_chunks,
                                    uint32_t *slot_count, uint32_t *quads, uint32_t per_slot,
                                    unsigned int *ticket, hipStream_t stream) {
  int per_cu = 0;
  MESO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, mesh_slotted_kernel<GREEDY, SRC>, kMeshBlock, 0));
  if (per_cu < 1) per_cu = 1;
  int grid = per_cu * ctx->props.multiProcessorCount;
  if (grid > max_chunks) grid = max_chunks;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((mesh_slotted_kernel<GREEDY, SRC>), dim3(grid), dim3(kMeshBlock), 0, stream,
                     scene->dev(), xyz, slots, n_dev, (uint32_t)max_chunks, slot_count, quads,
                     per_slot, ticket);
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

meso_status launch_mesh_slotted(meso_context *ctx, const meso_scene *scene, const int32_t *xyz,
                                const uint32_t *slots, const uint32_t *n_dev, int32_t max_chunks,
                                uint32_t mode, uint32_t *slot_count, uint32_t *quads,
                                uint32_t per_slot, unsigned int *ticket, hipStream_t stream) {
  const bool g = mode == MESO_MESH_GREEDY;
  if (scene->kind == MESO_SCENE_BRICKMAP_U16)
    return g ? launch_slotted_t<true, 2>(ctx, scene, xyz, slots, n_dev, max_chunks, slot_count,
                                         quads, per_slot, ticket, stream)
             : launch_slotted_t<false, 2>(ctx, scene, xyz, slots, n_dev, max_chunks, slot_count,
                                          quads, per_slot, ticket, stream);
  return g ? launch_slotted_t<true, 3>(ctx, scene, xyz, slots, n_dev, max_chunks, slot_count,
                                       quads, per_slot, ticket, stream)
           : launch_slotted_t<false, 3>(ctx, scene, xyz, slots, n_dev, max_chunks, slot_count,
                                        quads, per_slot, ticket, stream);
}

template <int ES, bool GREEDY>
static void launch_plain_t(const void *voxels, uint32_t *d_count, uint32_t *d_quads,
                           uint32_t capacity, hipStream_t stream) {
  hipLaunchKernelGGL((mesh_plain_kernel<ES, GREEDY>), dim3(1), dim3(kMeshBlock), 0, stream, voxels,
                     d_count, d_quads, capacity);
}

meso_status launch_mesh_plain(const void *voxels, int32_t es, uint32_t mode, uint32_t *d_count,
                              uint32_t *d_quads, uint32_t capacity, hipStream_t stream) {
  const bool g = mode == MESO_MESH_GREEDY;
  if (es == 1) {
    if (g) launch_plain_t<1, true>(voxels, d_count, d_quads, capacity, stream);
    else launch_plain_t<1, false>(voxels, d_count, d_quads, capacity, stream);
  } else {
    if (g) launch_plain_t<2, true>(voxels, d_count, d_quads, capacity, stream);
    else launch_plain_t<2, false>(voxels, d_count, d_quads, capacity, stream);
  }
  MESO_HIP(hipGetLastError());
  return MESO_OK;
}

}  // namespace meso

using namespace meso;

extern "C" {

uint64_t meso_mesh_workspace_bytes(int32_t n_chunks) {
  if (n_chunks < 1) n_chunks = 1;
  return 256u + 8u * (uint64_t)n_chunks;
}

meso_status meso_mesh_set_lookback_spin_limit(meso_context *ctx, uint32_t limit) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  ctx->lookback_spin_limit = limit;
  return MESO_OK;
}

meso_status meso_mesh_check(meso_context *ctx, const void *d_work, meso_stream stream) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  const void *w = d_work ? d_work : ctx->mesh_scratch;
  MESO_CHECK(w, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "no workspace");
  MESO_HIP(hipSetDevice(ctx->device));
  const hipStream_t s = stream_of(ctx, stream);
  MESO_HIP(hipStreamSynchronize(s));
  unsigned int err = 0;
  MESO_HIP(hipMemcpy(&err, (const uint8_t *)w + offsetof(MeshWork, error), sizeof(err),
                     hipMemcpyDeviceToHost));
  MESO_CHECK(err == 0, MESO_ERR_RUNTIME,
             "batched meshing: a chunk's look-back wait for its predecessors gave up after %u "
             "polls; its prefix, quads and the total are wrong",
             ctx->lookback_spin_limit);
  return MESO_OK;
}

meso_status meso_mesh_chunks_noalloc(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                     uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                     uint64_t capacity, uint64_t *d_total, meso_stream stream) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  MESO_CHECK(src && src->size >= sizeof(meso_chunk_source), MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "meso_chunk_source missing or too small");
  MESO_CHECK(src->n_chunks >= 1 && src->n_chunks <= MESO_MESH_NOALLOC_MAX_CHUNKS,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 1..%d: the static scratch holds the look-back state of at most "
             "%d chunks (meso_mesh_chunks takes any count with a workspace)",
             src->n_chunks, MESO_MESH_NOALLOC_MAX_CHUNKS, MESO_MESH_NOALLOC_MAX_CHUNKS);
  MESO_CHECK(ctx->mesh_scratch, MESO_ERR_RUNTIME, "mesher scratch address missing");
  // the scratch takes the place of the workspace; launch_mesh_batched clears
  // it with a fill kernel before the persistent launch
  return meso_mesh_chunks(ctx, src, mode, d_counts, d_prefix, d_quads, capacity, d_total,
                          ctx->mesh_scratch, stream);
}

meso_status meso_mesh_chunk(meso_context *ctx, const void *d_voxels, int32_t elem_bytes,
                            uint32_t mode, uint32_t *d_count, uint32_t *d_quads, uint32_t capacity,
                            meso_stream stream) {
  MESO_CHECK(ctx && d_voxels && d_count, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(((uintptr_t)d_voxels & 15u) == 0, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "chunk voxels must be 16-byte aligned");
  MESO_CHECK(capacity == 0 || d_quads, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "quads NULL");
  MESO_CHECK(elem_bytes == 1 || elem_bytes == 2, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "elem_bytes must be 1 or 2");
  MESO_CHECK(mode == MESO_MESH_CULLED || mode == MESO_MESH_GREEDY, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown mesh mode %u", mode);
  return launch_mesh_plain(d_voxels, elem_bytes, mode, d_count, d_quads, capacity,
                           stream_of(ctx, stream));
}

meso_status meso_mesh_scene_chunks(meso_context *ctx, meso_scene *scene, const int32_t *d_chunk_xyz,
                                   int32_t n_chunks, uint32_t mode, uint32_t *d_counts,
                                   uint32_t *d_prefix, uint32_t *d_quads, uint64_t capacity,
                                   uint64_t *d_total, void *d_work, meso_stream stream) {
  MESO_CHECK(ctx && scene && d_chunk_xyz, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(scene->ctx == ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "scene of another context");
  MESO_CHECK(scene->kind == MESO_SCENE_BRICKMAP_U16 || scene->kind == MESO_SCENE_TWO_LEVEL_U16,
             MESO_ERR_NOT_SUPPORTED, "scene meshing needs a brick map or two-level scene");
  MESO_CHECK(n_chunks >= 1 && n_chunks <= 40000, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "n_chunks %d not in 1..40000", n_chunks);
  MESO_CHECK(d_counts && d_prefix && d_total && d_work && (capacity == 0 || d_quads),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL output");
  MESO_CHECK(mode == MESO_MESH_CULLED || mode == MESO_MESH_GREEDY, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown mesh mode %u", mode);
  return launch_mesh_scene(ctx, scene, d_chunk_xyz, n_chunks, mode, d_counts, d_prefix, d_quads,
                           capacity, d_total, d_work, stream_of(ctx, stream));
}

meso_status meso_mesh_chunks(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                             uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                             uint64_t capacity, uint64_t *d_total, void *d_work,
                             meso_stream stream) {
  MESO_CHECK(ctx, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL context");
  meso_status st = check_source(src);
  if (st != MESO_OK) return st;
  MESO_CHECK(d_counts && d_prefix && d_total && d_work && (capacity == 0 || d_quads),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL output");
  MESO_CHECK(mode == MESO_MESH_CULLED || mode == MESO_MESH_GREEDY, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown mesh mode %u", mode);
  const MeshLaunch *shape = nullptr;
  MeshLaunch tuned;
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    char key[96];
    snprintf(key, sizeof(key), "mesh:%d:%d:%d:%u", src->elem_bytes, src->layout, src->n_chunks,
             mode);
    auto it = ctx->tune.find(key);
    if (it != ctx->tune.end()) {
      tuned = MeshLaunch{it->second.block, it->second.grid};
      shape = &tuned;
    }
  }
  return launch_mesh_batched(ctx, src, mode, d_counts, d_prefix, d_quads, capacity, d_total,
                             d_work, stream_of(ctx, stream), shape);
}

}  // extern "C"
