"""ctypes binding of the CPU oracle (oracle/meso_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product library
(mesoengine_amd) never imports this package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

EMPTY_SLOT = 0xFFFFFFFF
FACE_MISS = 0xFF
FACE_INSIDE = 6
SCENE_DENSE_U8, SCENE_BRICKMAP_U16, SCENE_TWO_LEVEL_U16 = 1, 2, 3
GEN_HEIGHTFIELD, GEN_TERRAIN_CAVES, GEN_CAVES = 1, 2, 3
RAY_SHADOW, RAY_AO = 1, 2


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("fwd", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("sx", C.c_float), ("sy", C.c_float),
                ("inv_w2", C.c_float), ("inv_h2", C.c_float),
                ("width", C.c_int32), ("height", C.c_int32)]


class Scene(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n", C.c_int32), ("dense", C.c_void_p),
                ("slots", C.c_void_p), ("masks", C.c_void_p), ("voxels", C.c_void_p),
                ("top", C.c_void_p), ("occ32", C.c_void_p), ("occ128", C.c_void_p),
                ("groups", C.c_void_p), ("box_lo", C.c_int32 * 3), ("box_hi", C.c_int32 * 3)]


class Counters(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("n_vox", "n_top", "n_node", "n_brickslot",
                                          "n_brickmask", "hits", "rays", "shadow_rays",
                                          "ao_rays", "n_coarse", "n_iter")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Hit(C.Structure):
    _fields_ = [("index", C.c_int64), ("face", C.c_int32), ("t", C.c_float),
                ("cell", C.c_int32 * 3), ("material", C.c_uint16)]


class Shading(C.Structure):
    _fields_ = [("light", C.c_float * 3), ("ao_seed", C.c_uint32), ("ao_samples", C.c_int32),
                ("ao_radius", C.c_float)]


class PoolParams(C.Structure):
    _fields_ = [("forward_radius", C.c_int32), ("backward_radius", C.c_int32),
                ("view_threshold", C.c_float), ("capacity", C.c_int32),
                ("quads_per_slot", C.c_int32), ("max_loads", C.c_int32), ("greedy", C.c_int32)]


class PoolResult(C.Structure):
    _fields_ = [("n_candidates", C.c_uint32), ("n_loads", C.c_uint32),
                ("n_evicted", C.c_uint32)]


class ChunkParams(C.Structure):
    _fields_ = [("fill", C.c_float), ("noise_seed", C.c_uint32), ("n_spheres", C.c_int32),
                ("sph", (C.c_int32 * 4) * 4)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        P = C.c_void_p
        sig = {
            "mo_quad_pack": (C.c_uint32, [C.c_int] * 6),
            "mo_splitmix64": (C.c_uint64, [P]),
            "mo_pcg32_srandom": (None, [P, C.c_uint64, C.c_uint64]),
            "mo_pcg32_next": (C.c_uint32, [P]),
            "mo_pcg32_seed": (None, [P, C.c_uint64]),
            "mo_hash32": (C.c_uint32, [C.c_uint32]),
            "mo_lattice_hash3": (C.c_uint32, [C.c_int32] * 3 + [C.c_uint32]),
            "mo_value_noise3": (C.c_float, [C.c_float] * 3 + [C.c_uint32]),
            "mo_value_noise2": (C.c_float, [C.c_float] * 2 + [C.c_uint32]),
            "mo_fbm2": (C.c_float, [C.c_float, C.c_float, C.c_int, C.c_float, C.c_uint32]),
            "mo_fbm3": (C.c_float, [C.c_float] * 3 + [C.c_int, C.c_float, C.c_uint32]),
            "mo_sphere_params": (None, [C.c_uint64, C.c_int, C.c_int, P]),
            "mo_gen_spheres_dense": (None, [P, C.c_int, C.c_int, P]),
            "mo_height": (C.c_int32, [C.c_int32] * 3 + [C.c_uint32]),
            "mo_scene_voxel": (C.c_uint16, [C.c_int] + [C.c_int32] * 4 + [C.c_uint32]),
            "mo_build_brickmap": (C.c_int64, [C.c_int, C.c_int32, C.c_uint32, P, P, P, C.c_int64]),
            "mo_brickmap_from_dense": (C.c_int64, [P, C.c_int32, P, P, P, C.c_int64]),
            "mo_gen_brick": (C.c_int32, [C.c_int, C.c_int32, C.c_uint32] + [C.c_int32] * 3 + [P, P]),
            "mo_chunk_params_gen": (None, [C.c_uint64, C.c_int32, P]),
            "mo_gen_chunk": (None, [P, P]),
            "mo_camera_derive": (None, [P, P, P, C.c_float, C.c_int32, C.c_int32, P]),
            "mo_primary_ray": (None, [P, C.c_int32, C.c_int32, P, P]),
            "mo_trace": (C.c_int, [P, P, P, P, P]),
            "mo_occluded": (C.c_int, [P, P, P, P, C.c_float, P]),
            "mo_raymarch": (None, [P, P, C.c_uint32, P, C.c_int32, C.c_int32, P, P, P, P, P]),
            "mo_raymarch_mt": (None, [P, P, C.c_uint32, P, C.c_int32, C.c_int32, C.c_int32,
                                      P, P, P, P, P]),
            "mo_ao_direction": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]),
            "mo_mesh_chunk": (C.c_uint32, [P, C.c_int, C.c_int, P, C.c_uint32]),
            "mo_mesh_chunks_mt": (C.c_uint64, [P, C.c_int, C.c_int32, C.c_int, C.c_int32, P]),
            "mo_ref_erode_mips": (None, [P, C.c_int32, C.c_int32, P]),
            "mo_ref_generate_sphere": (C.c_int32, [C.c_int32] * 3 + [C.c_float, C.c_int32, P]),
            "mo_ref_test_generator": (C.c_int32, [C.c_int32] * 3 + [C.c_float, C.c_int32, P]),
            "mo_ref_hash3": (C.c_double, [C.c_double] * 3),
            "mo_ref_noised": (None, [P, P]),
            "mo_ref_displacement": (C.c_double, [P]),
            "mo_ref_octant_id": (C.c_int32, [C.c_float] * 3),
            "mo_build_coarse": (None, [P, C.c_int64, P, P, P]),
            "mo_edit_chunks": (None, [P, P, P, P, C.c_int32, C.c_uint32]),
            "mo_scene_chunk": (None, [P, C.c_int32, C.c_int32, C.c_int32, P]),
            "mo_scene_box": (None, [P, P, P]),
            "mo_pool_view_threshold": (C.c_float, [C.c_float]),
            "mo_pool_forward": (None, [P, P]),
            "mo_pool_bake_directions": (None, [C.c_int32, P]),
            "mo_pool_snap": (None, [P, C.c_int32, P, P]),
            "mo_chunk_importance": (C.c_float, [C.c_int32] * 3 + [P]),
            "mo_chunk_desired": (C.c_int, [C.c_int32] * 3 + [P, P]),
            "mo_chunk_nonempty": (C.c_int, [P] + [C.c_int32] * 3),
            "mo_pool_update": (None, [P, P, C.c_uint32, P, P, P, P, P, P, P, P, P, P]),
            "mo_raster_pool": (None, [P, C.c_int32, C.c_int32, P, P, P, C.c_int32, C.c_int32,
                                      P, P, P, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------- RNG ----------------
class PCG32(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


def pcg32_srandom(initstate, initseq):
    r = PCG32()
    lib().mo_pcg32_srandom(C.byref(r), initstate, initseq)
    return r


def pcg32_next(r):
    return lib().mo_pcg32_next(C.byref(r))


def splitmix64_seq(seed, n):
    s = C.c_uint64(seed)
    return [lib().mo_splitmix64(C.byref(s)) for _ in range(n)]


# ---------------- generators ----------------
def sphere_params(seed, n, n_spheres):
    p = np.zeros(4 * n_spheres, np.int32)
    lib().mo_sphere_params(seed, n, n_spheres, _p(p))
    return p.reshape(n_spheres, 4)


def spheres_dense(seed, n=128, n_spheres=8):
    p = sphere_params(seed, n, n_spheres)
    g = np.zeros(n ** 3, np.uint8)
    lib().mo_gen_spheres_dense(_p(np.ascontiguousarray(p.ravel())), n_spheres, n, _p(g))
    return g.reshape(n, n, n)  # [z][y][x]


def build_brickmap(kind, n, seed):
    nb = n // 8
    slots = np.zeros(nb ** 3, np.uint32)
    cnt = lib().mo_build_brickmap(kind, n, seed, _p(slots), None, None, 0)
    masks = np.zeros((max(cnt, 1), 8), np.uint64)
    vox = np.zeros((max(cnt, 1), 512), np.uint16)
    cnt2 = lib().mo_build_brickmap(kind, n, seed, _p(slots), _p(masks), _p(vox), cnt)
    assert cnt2 == cnt
    return slots, masks[:cnt], vox[:cnt]


def brickmap_from_dense(grid_u16):
    n = grid_u16.shape[0]
    nb = n // 8
    g = np.ascontiguousarray(grid_u16, dtype=np.uint16)
    slots = np.zeros(nb ** 3, np.uint32)
    cnt = lib().mo_brickmap_from_dense(_p(g), n, _p(slots), None, None, 0)
    masks = np.zeros((max(cnt, 1), 8), np.uint64)
    vox = np.zeros((max(cnt, 1), 512), np.uint16)
    lib().mo_brickmap_from_dense(_p(g), n, _p(slots), _p(masks), _p(vox), cnt)
    return slots, masks[:cnt], vox[:cnt]


def gen_brick(kind, n, seed, bx, by, bz):
    vox = np.zeros(512, np.uint16)
    mask = np.zeros(8, np.uint64)
    s = lib().mo_gen_brick(kind, n, seed, bx, by, bz, _p(vox), _p(mask))
    return s, vox, mask


def chunk_params(seed, n_chunks):
    arr = (ChunkParams * n_chunks)()
    lib().mo_chunk_params_gen(seed, n_chunks, C.cast(arr, C.c_void_p))
    return arr


def gen_chunks(seed, n_chunks):
    params = chunk_params(seed, n_chunks)
    out = np.zeros((n_chunks, 32768), np.uint16)
    for c in range(n_chunks):
        lib().mo_gen_chunk(C.cast(C.byref(params[c]), C.c_void_p), _p(out[c]))
    return out, params


# ---------------- camera / traversal ----------------
def camera(eye, target, up, vfov_deg, width, height):
    cam = Camera()
    f3 = C.c_float * 3
    lib().mo_camera_derive(C.cast(f3(*eye), C.c_void_p), C.cast(f3(*target), C.c_void_p),
                           C.cast(f3(*up), C.c_void_p), vfov_deg, width, height,
                           C.cast(C.byref(cam), C.c_void_p))
    return cam


def primary_ray(cam, px, py):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().mo_primary_ray(C.cast(C.byref(cam), C.c_void_p), px, py, _p(o), _p(d))
    return o, d


class OracleScene:
    """Keeps the numpy arrays alive behind an mo_scene struct; builds the
    empty-space levels (occ32 / occ128 or node groups) of brick scenes."""

    def __init__(self, kind, n, dense=None, slots=None, masks=None, voxels=None, top=None):
        self.arrays = [a for a in (dense, slots, masks, voxels, top) if a is not None]
        self.s = Scene(kind, n, _p(dense), _p(slots), _p(masks), _p(voxels), _p(top),
                       None, None, None)
        self.kind, self.n = kind, n
        self.occ32 = self.occ128 = self.groups = None
        if kind == SCENE_BRICKMAP_U16:
            n32, n128 = -(-n // 32), -(-n // 128)
            self.occ32 = np.zeros((n32 ** 3 + 31) // 32, np.uint32)
            self.occ128 = np.zeros((n128 ** 3 + 31) // 32, np.uint32)
            self.s.occ32, self.s.occ128 = _p(self.occ32), _p(self.occ128)
            lib().mo_build_coarse(self.ptr, 0, _p(self.occ32), _p(self.occ128), None)
        elif kind == SCENE_TWO_LEVEL_U16:
            n_nodes = len(slots) // 4096
            self.groups = np.zeros(max(n_nodes, 1), np.uint64)
            self.s.groups = _p(self.groups)
            lib().mo_build_coarse(self.ptr, n_nodes, None, None, _p(self.groups))
        lib().mo_scene_box(self.ptr, self.s.box_lo, self.s.box_hi)

    @property
    def box(self):
        return tuple(self.s.box_lo), tuple(self.s.box_hi)

    @property
    def ptr(self):
        return C.cast(C.byref(self.s), C.c_void_p)


def dense_scene(grid_u8):
    g = np.ascontiguousarray(grid_u8, dtype=np.uint8).ravel()
    return OracleScene(SCENE_DENSE_U8, int(round(g.size ** (1 / 3))), dense=g)


def brickmap_scene(n, slots, masks, voxels):
    return OracleScene(SCENE_BRICKMAP_U16, n, slots=np.ascontiguousarray(slots),
                       masks=np.ascontiguousarray(masks), voxels=np.ascontiguousarray(voxels))


def two_level_scene(n, top, node_slots, masks, voxels):
    return OracleScene(SCENE_TWO_LEVEL_U16, n, top=np.ascontiguousarray(top),
                       slots=np.ascontiguousarray(node_slots), masks=np.ascontiguousarray(masks),
                       voxels=np.ascontiguousarray(voxels))


def trace(scene, o, d):
    h = Hit()
    cnt = Counters()
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    lib().mo_trace(scene.ptr, _p(o), _p(d), C.cast(C.byref(h), C.c_void_p),
                   C.cast(C.byref(cnt), C.c_void_p))
    return h, cnt


def occluded(scene, o, d, start, tmax=float("inf")):
    cnt = Counters()
    o = np.asarray(o, np.float32)
    d = np.asarray(d, np.float32)
    st = np.asarray(start, np.int32)
    r = lib().mo_occluded(scene.ptr, _p(o), _p(d), _p(st), tmax, C.cast(C.byref(cnt), C.c_void_p))
    return bool(r), cnt


def shading(light=(0.4, 1.0, 0.3), ao_seed=0, ao_samples=4, ao_radius=16.0):
    lx, ly, lz = (np.float32(v) for v in light)
    ln = np.sqrt(np.float32(np.float32(lx * lx) + np.float32(ly * ly)) + np.float32(lz * lz),
                 dtype=np.float32)
    L = (C.c_float * 3)(float(lx / ln), float(ly / ln), float(lz / ln))
    return Shading(L, ao_seed, ao_samples, ao_radius)


def raymarch(scene, cam, flags, sh, row0=0, row1=None, threads=1):
    W, H = cam.width, cam.height
    row1 = H if row1 is None else row1
    n = W * H
    idx = np.full(n, -1, np.int64)
    face = np.full(n, FACE_MISS, np.uint8)
    t = np.full(n, np.inf, np.float32)
    shade = np.zeros(n, np.float32)
    cnt = Counters()
    V = C.c_void_p
    if threads > 1:
        lib().mo_raymarch_mt(scene.ptr, C.cast(C.byref(cam), V), flags, C.cast(C.byref(sh), V),
                             row0, row1, threads, _p(idx), _p(face), _p(t), _p(shade),
                             C.cast(C.byref(cnt), V))
    else:
        lib().mo_raymarch(scene.ptr, C.cast(C.byref(cam), V), flags, C.cast(C.byref(sh), V),
                          row0, row1, _p(idx), _p(face), _p(t), _p(shade), C.cast(C.byref(cnt), V))
    return idx, face, t, shade, cnt


def ao_direction(seed, pixel, sample, face):
    d = np.zeros(3, np.float32)
    lib().mo_ao_direction(seed, pixel, sample, face, _p(d))
    return d


# ---------------- meshing ----------------
def mesh_chunk(vox, greedy):
    v = np.ascontiguousarray(vox)
    es = v.dtype.itemsize
    cnt = lib().mo_mesh_chunk(_p(v), es, int(greedy), None, 0)
    q = np.zeros(max(cnt, 1), np.uint32)
    lib().mo_mesh_chunk(_p(v), es, int(greedy), _p(q), cnt)
    return q[:cnt]


def mesh_chunks_counts(vox, greedy, threads=1):
    v = np.ascontiguousarray(vox)
    n_chunks = v.shape[0]
    counts = np.zeros(n_chunks, np.uint32)
    total = lib().mo_mesh_chunks_mt(_p(v), v.dtype.itemsize, n_chunks, int(greedy), threads,
                                    _p(counts))
    return counts, total


def unpack_quads(q):
    q = np.asarray(q, np.uint32)
    return {"x": q & 31, "y": (q >> 5) & 31, "z": (q >> 10) & 31, "w": ((q >> 15) & 31) + 1,
            "h": ((q >> 20) & 31) + 1, "face": (q >> 25) & 7, "pad": q >> 28}


# ---------------- reference restatements ----------------
def ref_erode_mips(occ, depth=4):
    occ = np.ascontiguousarray(occ, dtype=np.uint8)
    res = occ.shape[0]
    mips = np.zeros((depth,) + occ.shape, np.uint8)
    lib().mo_ref_erode_mips(_p(occ), res, depth, _p(mips))
    return mips


def ref_generate_sphere(cx, cy, cz, block_size=1.0, res=16):
    cells = np.zeros((res ** 3, 3), np.uint8)
    k = lib().mo_ref_generate_sphere(cx, cy, cz, block_size, res, _p(cells))
    return cells[:k]


def ref_test_generator(cx, cy, cz, block_size=1.0, res=16):
    cells = np.zeros((res ** 3, 3), np.uint8)
    k = lib().mo_ref_test_generator(cx, cy, cz, block_size, res, _p(cells))
    return cells[:k]


def ref_hash3(x, y, z):
    return lib().mo_ref_hash3(float(x), float(y), float(z))


def ref_noised(x):
    """(dx, dy, dz, value) of the reference's gradient value noise"""
    xi = np.ascontiguousarray(x, np.float64)
    out = np.zeros(4, np.float64)
    lib().mo_ref_noised(_p(xi), _p(out))
    return out


def ref_displacement(p):
    return lib().mo_ref_displacement(_p(np.ascontiguousarray(p, np.float64)))


def two_level_from_bricks(n, slots, masks, voxels):
    """Two-level map (top node = 16^3 bricks) from a brick map of edge n
    (multiple of 128).  Nodes are numbered in linear top order; bricks are
    renumbered node-major: node by node, local index bx + 16 by + 256 bz."""
    nb, nt = n // 8, n // 128
    s = np.asarray(slots, np.uint32).reshape(nb, nb, nb)  # [bz][by][bx]
    top = np.full(nt ** 3, EMPTY_SLOT, np.uint32)
    node_slots, new_masks, new_vox = [], [], []
    nid, bid = 0, 0
    for tz in range(nt):
        for ty in range(nt):
            for tx in range(nt):
                blk = s[tz * 16:(tz + 1) * 16, ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16]
                if (blk == EMPTY_SLOT).all():
                    continue
                top[tx + nt * ty + nt * nt * tz] = nid
                nid += 1
                flat = blk.reshape(-1)  # local = bx + 16 by + 256 bz
                tbl = np.full(4096, EMPTY_SLOT, np.uint32)
                used = np.nonzero(flat != EMPTY_SLOT)[0]
                tbl[used] = np.arange(bid, bid + len(used), dtype=np.uint32)
                bid += len(used)
                new_masks.append(masks[flat[used]])
                new_vox.append(voxels[flat[used]])
                node_slots.append(tbl)
    ns = np.concatenate(node_slots) if node_slots else np.zeros(0, np.uint32)
    m = np.concatenate(new_masks) if new_masks else np.zeros((0, 8), np.uint64)
    v = np.concatenate(new_vox) if new_vox else np.zeros((0, 512), np.uint16)
    return top, ns, m, v


def dense_from_bricks(n, slots, voxels):
    """Dense u16 grid [z][y][x] from a brick map."""
    nb = n // 8
    g = np.zeros((n, n, n), np.uint16)
    s = np.asarray(slots, np.uint32).reshape(nb, nb, nb)
    for bz, by, bx in zip(*np.nonzero(s != EMPTY_SLOT)):
        g[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = \
            voxels[s[bz, by, bx]].reshape(8, 8, 8)
    return g


def edit_chunks(scene, chunk_xyz, seed):
    """Applies the seeded chunk edits to the scene's voxel and mask arrays in place."""
    xyz = np.ascontiguousarray(chunk_xyz, np.int32).ravel()
    lib().mo_edit_chunks(scene.ptr, scene.s.voxels, scene.s.masks, _p(xyz), len(xyz) // 3, seed)


def scene_chunk(scene, cx, cy, cz):
    out = np.zeros(32768, np.uint16)
    lib().mo_scene_chunk(scene.ptr, cx, cy, cz, _p(out))
    return out


# ---------------- chunk pool ----------------
NO_SLOT = 0xFFFFFFFF


def pool_view_threshold(view_angle_deg):
    return lib().mo_pool_view_threshold(view_angle_deg)


def pool_forward(raw):
    r = np.ascontiguousarray(raw, np.float32)
    f = np.zeros(3, np.float32)
    lib().mo_pool_forward(_p(r), _p(f))
    return f


def pool_bake_directions(n):
    d = np.zeros(3 * n, np.float32)
    lib().mo_pool_bake_directions(n, _p(d))
    return d.reshape(n, 3)


def pool_snap(dirs, f):
    d = np.ascontiguousarray(dirs, np.float32)
    fv = np.ascontiguousarray(f, np.float32)
    out = np.zeros(3, np.float32)
    lib().mo_pool_snap(_p(d), len(d), _p(fv), _p(out))
    return out


def chunk_importance(dx, dy, dz, f):
    f = np.ascontiguousarray(f, np.float32)
    return lib().mo_chunk_importance(dx, dy, dz, _p(f))


def pool_params(forward_radius=24, backward_radius=6, view_angle=120.0, capacity=16384,
                quads_per_slot=8192, max_loads=256, greedy=True):
    return PoolParams(forward_radius, backward_radius, pool_view_threshold(view_angle), capacity,
                      quads_per_slot, max_loads, 1 if greedy else 0)


def chunk_desired(dx, dy, dz, f, params):
    f = np.ascontiguousarray(f, np.float32)
    return bool(lib().mo_chunk_desired(dx, dy, dz, _p(f), C.byref(params)))


def chunk_nonempty(scene, cx, cy, cz):
    return bool(lib().mo_chunk_nonempty(scene.ptr, cx, cy, cz))


class OraclePool:
    """The chunk pool's state (lookup table, slots, quads) and its updates."""

    def __init__(self, scene, params, bake_directions=0):
        """bake_directions > 0: the view is the nearest of that many baked
        directions (the reference's BakeVisibilityViewNum)"""
        self.scene, self.p = scene, params
        self.dirs = pool_bake_directions(bake_directions) if bake_directions else None
        self.nc = -(-scene.n // 32)
        self.lookup = np.full(self.nc ** 3, NO_SLOT, np.uint32)
        cap = params.capacity
        self.slot_chunk = np.full(cap, NO_SLOT, np.uint32)
        self.slot_stamp = np.zeros(cap, np.uint32)
        self.slot_count = np.zeros(cap, np.uint32)
        self.quads = np.zeros(cap * params.quads_per_slot, np.uint32)
        self.load_xyz = np.zeros(3 * max(params.max_loads, 1), np.int32)
        self.load_slot = np.zeros(max(params.max_loads, 1), np.uint32)
        self.stamp = 0

    def update(self, cam_chunk, forward_raw):
        """one update; returns (n_candidates, n_loads, n_evicted, loads [(x, y, z, slot)])"""
        self.stamp = (self.stamp + 1) & 0xFFFFFFFF
        f = pool_forward(forward_raw)
        if self.dirs is not None:
            f = pool_snap(self.dirs, f)
        cam = np.ascontiguousarray(cam_chunk, np.int32)
        r = PoolResult()
        lib().mo_pool_update(self.scene.ptr, C.byref(self.p), self.stamp, _p(cam), _p(f),
                             _p(self.lookup), _p(self.slot_chunk), _p(self.slot_stamp),
                             _p(self.slot_count), _p(self.quads), _p(self.load_xyz),
                             _p(self.load_slot), C.byref(r))
        n = r.n_loads
        loads = [tuple(int(v) for v in self.load_xyz[3 * i:3 * i + 3]) + (int(self.load_slot[i]),)
                 for i in range(n)]
        return r.n_candidates, n, r.n_evicted, loads

    def remesh(self, xyz):
        """re-meshes the resident chunks of a dirty list into their slots"""
        q, nc = self.p.quads_per_slot, self.nc
        for x, y, z in np.asarray(xyz, np.int64).reshape(-1, 3):
            if min(x, y, z) < 0 or max(x, y, z) >= nc:
                continue
            sl = int(self.lookup[x + nc * (y + nc * z)])
            if sl == NO_SLOT:
                continue
            ref = mesh_chunk(scene_chunk(self.scene, int(x), int(y), int(z)), bool(self.p.greedy))
            self.slot_count[sl] = len(ref)
            k = min(len(ref), q)
            self.quads[sl * q: sl * q + k] = ref[:k]

    def slot_quads(self, slot):
        q = self.p.quads_per_slot
        k = min(int(self.slot_count[slot]), q)
        return self.quads[slot * q: slot * q + k]


# ---------------- quad visibility ----------------
def raster_pool(cam, n, nc, slot_chunk, slot_count, quads, per_slot, light):
    """(index, face, t, shade, keys) of the nearest front-facing quad per pixel"""
    w, h = cam.width, cam.height
    keys = np.empty(w * h, np.uint64)
    idx = np.empty(w * h, np.int64)
    face = np.empty(w * h, np.uint8)
    t = np.empty(w * h, np.float32)
    shade = np.empty(w * h, np.float32)
    sc = np.ascontiguousarray(slot_chunk, np.uint32)
    cnt = np.ascontiguousarray(slot_count, np.uint32)
    q = np.ascontiguousarray(quads, np.uint32)
    li = np.ascontiguousarray(light, np.float32)
    lib().mo_raster_pool(C.byref(cam), n, nc, _p(sc), _p(cnt), _p(q), len(sc), per_slot, _p(li),
                         _p(keys), _p(idx), _p(face), _p(t), _p(shade))
    return idx, face, t, shade, keys
