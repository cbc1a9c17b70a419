"""Python harness over the C-ABI: device buffers, scenes, raymarch, meshing,
serving.  Every call goes through libmeso_hip.so; there is no CPU fallback."""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import byref, check


class Context:
    def __init__(self, device=0):
        self.lib = L.load()
        h = C.c_void_p()
        check(self.lib.meso_context_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.meso_context_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_info(self):
        d = L.DeviceInfo()
        d.size = C.sizeof(d)
        check(self.lib.meso_device_info_get(self.h, byref(d)))
        return d

    def sync(self, stream=None):
        check(self.lib.meso_stream_sync(self.h, stream))

    # ---------------- buffers
    def buffer(self, nbytes, init=None):
        return Buffer(self, nbytes, init)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        return Buffer(self, max(a.nbytes, 1), a if a.nbytes else None)

    def zeros(self, nbytes, stream=None):
        b = Buffer(self, max(int(nbytes), 16))
        check(self.lib.meso_buffer_fill_zero(b.h, stream))
        return b

    # ---------------- scenes
    def scene_generate(self, generator, layout, n, seed, n_spheres=8):
        d = L.GenerateDesc(C.sizeof(L.GenerateDesc), generator, layout, n, seed, n_spheres)
        h = C.c_void_p()
        check(self.lib.meso_scene_generate(self.h, byref(d), C.byref(h)))
        return Scene(self, h)

    def scene_dense(self, grid_u8):
        g = np.ascontiguousarray(grid_u8, dtype=np.uint8)
        n = g.shape[0]
        h = C.c_void_p()
        check(self.lib.meso_scene_create_dense_u8(self.h, n, g.ctypes.data, C.byref(h)))
        return Scene(self, h)

    def scene_brickmap(self, n, slots, masks, voxels):
        s = np.ascontiguousarray(slots, np.uint32)
        m = np.ascontiguousarray(masks, np.uint64)
        v = np.ascontiguousarray(voxels, np.uint16)
        h = C.c_void_p()
        check(self.lib.meso_scene_create_brickmap_u16(self.h, n, s.ctypes.data, len(m),
                                                      m.ctypes.data if len(m) else None,
                                                      v.ctypes.data if len(v) else None,
                                                      C.byref(h)))
        return Scene(self, h)

    def scene_two_level(self, n, top, node_slots, masks, voxels):
        t = np.ascontiguousarray(top, np.uint32)
        ns = np.ascontiguousarray(node_slots, np.uint32)
        m = np.ascontiguousarray(masks, np.uint64)
        v = np.ascontiguousarray(voxels, np.uint16)
        h = C.c_void_p()
        check(self.lib.meso_scene_create_two_level_u16(
            self.h, n, t.ctypes.data, len(ns) // 4096, ns.ctypes.data if len(ns) else None,
            len(m), m.ctypes.data if len(m) else None, v.ctypes.data if len(v) else None,
            C.byref(h)))
        return Scene(self, h)


class Buffer:
    def __init__(self, ctx, nbytes, init=None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.nbytes = int(nbytes)
        h = C.c_void_p()
        src = None
        if init is not None:
            init = np.ascontiguousarray(init)
            assert init.nbytes == self.nbytes
            src = init.ctypes.data
        check(self.lib.meso_buffer_create(ctx.h, self.nbytes, src, C.byref(h)))
        self.h = h
        self.ptr = self.lib.meso_buffer_device_address(h, 0)
        if not self.ptr:
            raise L.MesoError(L.ERR_RUNTIME, "no device address")

    def download(self, dtype, count=None, offset=0):
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        out = np.empty(count, dt)
        if count:
            check(self.lib.meso_buffer_download(self.h, out.ctypes.data, out.nbytes, offset))
        return out

    def upload(self, arr, offset=0):
        a = np.ascontiguousarray(arr)
        check(self.lib.meso_buffer_upload(self.h, a.ctypes.data, a.nbytes, offset))

    def zero(self, stream=None):
        check(self.lib.meso_buffer_fill_zero(self.h, stream))

    def fill(self, byte):
        """every byte set to `byte`, by a synchronous upload"""
        self.upload(np.full(self.nbytes, byte, np.uint8))

    def free(self):
        if self.h:
            self.lib.meso_buffer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scene:
    def __init__(self, ctx, h):
        self.ctx = ctx
        self.h = h
        self.lib = ctx.lib

    def info(self):
        i = L.SceneInfo()
        i.size = C.sizeof(i)
        check(self.lib.meso_scene_get_info(self.h, byref(i)))
        return i

    def download(self, part, dtype, count):
        out = np.empty(count, dtype)
        if count:
            check(self.lib.meso_scene_download(self.h, part, out.ctypes.data, out.nbytes))
        return out

    def arrays(self):
        """Download the scene's arrays (for parity checks)."""
        i = self.info()
        n = i.n
        if i.kind == L.SCENE_DENSE_U8:
            return {"dense": self.download(L.PART_DENSE, np.uint8, n ** 3)}
        out = {"masks": self.download(L.PART_MASKS, np.uint64, i.n_bricks * 8).reshape(-1, 8),
               "voxels": self.download(L.PART_VOXELS, np.uint16, i.n_bricks * 512).reshape(-1, 512)}
        if i.kind == L.SCENE_BRICKMAP_U16:
            out["slots"] = self.download(L.PART_SLOTS, np.uint32, (n // 8) ** 3)
        else:
            out["top"] = self.download(L.PART_TOP, np.uint32, (n // 128) ** 3)
            out["slots"] = self.download(L.PART_SLOTS, np.uint32, i.n_nodes * 4096)
        return out

    def destroy(self):
        if self.h:
            self.lib.meso_scene_destroy(self.h)
            self.h = None


def camera(eye, target, up, vfov_deg, width, height):
    c = L.Camera()
    c.size = C.sizeof(c)
    c.eye[:] = [float(v) for v in eye]
    c.target[:] = [float(v) for v in target]
    c.up[:] = [float(v) for v in up]
    c.vfov_deg = vfov_deg
    c.width, c.height = width, height
    return c


def shading(light=(0.4, 1.0, 0.3), ao_seed=0, ao_samples=4, ao_radius=16.0):
    s = L.Shading()
    s.size = C.sizeof(s)
    s.light[:] = [float(v) for v in light]
    s.ao_seed, s.ao_samples, s.ao_radius = ao_seed, ao_samples, ao_radius
    return s


class RayTargets:
    """Device output buffers for one image."""

    def __init__(self, ctx, width, height, index_bytes=4, counters=False):
        self.n = width * height
        self.index_bytes = index_bytes
        self.index = ctx.buffer(self.n * index_bytes)
        self.face = ctx.buffer(self.n)
        self.t = ctx.buffer(self.n * 4)
        self.shade = ctx.buffer(self.n * 4)
        self.counters = ctx.zeros(8 * len(L.CNT_NAMES)) if counters else None
        o = L.RayOutputs()
        o.size = C.sizeof(o)
        o.hit_index, o.index_bytes = self.index.ptr, index_bytes
        o.face, o.t, o.shade = self.face.ptr, self.t.ptr, self.shade.ptr
        o.counters = self.counters.ptr if counters else None
        self.out = o

    def results(self):
        idx = self.index.download(np.int64 if self.index_bytes == 8 else np.int32)
        return (idx.astype(np.int64), self.face.download(np.uint8), self.t.download(np.float32),
                self.shade.download(np.float32))

    def counts(self):
        c = self.counters.download(np.uint64)
        return dict(zip(L.CNT_NAMES, (int(v) for v in c)))


def raymarch(ctx, scene, cam, flags, sh, targets, stream=None):
    check(ctx.lib.meso_raymarch(ctx.h, scene.h, byref(cam), flags, byref(sh) if sh else None,
                                byref(targets.out), stream))


def matches_tiled_frame(ctx, scene, cam, flags, sh, targets, shape):
    """One frame with the launch shape `shape` (a TuneResult) pinned, against
    one with 16x16 tiles (variant 0).  Returns True when index, face, t and
    shade are the same bits; `shape` stays pinned afterwards."""
    def frame(pin):
        check(ctx.lib.meso_autotune_pin_raymarch(ctx.h, scene.h, byref(cam), flags, byref(pin)))
        raymarch(ctx, scene, cam, flags, sh, targets)
        ctx.sync()
        return targets.results()

    tiled = L.TuneResult()
    tiled.size = C.sizeof(tiled)
    tiled.variant, tiled.grid, tiled.tile_w, tiled.tile_h = 0, 0, 16, 16
    ref = frame(tiled)
    got = frame(shape)
    return all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(ref, got))


def chunk_source(buf, elem_bytes, n_chunks, layout=L.CHUNKS_PACKED, grid_n=0):
    s = L.ChunkSource()
    s.size = C.sizeof(s)
    s.voxels = buf.ptr
    s.elem_bytes, s.layout, s.grid_n, s.n_chunks = elem_bytes, layout, grid_n, n_chunks
    return s


class MeshTargets:
    def __init__(self, ctx, n_chunks, capacity):
        self.n_chunks = n_chunks
        self.capacity = int(capacity)
        self.counts = ctx.buffer(4 * n_chunks)
        self.prefix = ctx.buffer(4 * n_chunks)
        self.quads = ctx.buffer(4 * max(self.capacity, 1))
        self.total = ctx.buffer(16)
        self.work = ctx.buffer(ctx.lib.meso_mesh_workspace_bytes(n_chunks))

    def results(self):
        counts = self.counts.download(np.uint32)
        prefix = self.prefix.download(np.uint32)
        total = int(self.total.download(np.uint64, 1)[0])
        quads = self.quads.download(np.uint32, min(total, self.capacity))
        return counts, prefix, total, quads


def mesh_chunks(ctx, src, mode, mt, stream=None):
    check(ctx.lib.meso_mesh_chunks(ctx.h, byref(src), mode, mt.counts.ptr, mt.prefix.ptr,
                                   mt.quads.ptr, mt.capacity, mt.total.ptr, mt.work.ptr, stream))


def mesh_chunk(ctx, vox_ptr, elem_bytes, mode, count_ptr, quads_ptr, capacity, stream=None):
    check(ctx.lib.meso_mesh_chunk(ctx.h, vox_ptr, elem_bytes, mode, count_ptr, quads_ptr,
                                  capacity, stream))


class Event:
    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.meso_event_create(ctx.h, C.byref(h)))
        self.h = h

    def record(self, stream=None):
        check(self.ctx.lib.meso_event_record(self.h, stream))

    def sync(self):
        check(self.ctx.lib.meso_event_sync(self.h))

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(self.ctx.lib.meso_event_elapsed_ms(self.h, end.h, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            self.ctx.lib.meso_event_destroy(self.h)
        except Exception:
            pass


class Server:
    """The continuous-batching server: requests are admitted by the HBM
    budget into `max_inflight` ring entries that one persistent kernel of
    `workgroups_per_cu` workgroups of 4 waves per CU (0: all that fit) draws
    8x8-pixel packets from."""

    def __init__(self, ctx, budget_bytes, max_pending=256, max_inflight=64, workgroups_per_cu=0,
                 idle_exit_us=0):
        self.ctx = ctx
        cfg = L.ServeConfig(C.sizeof(L.ServeConfig), budget_bytes, max_pending, max_inflight,
                            workgroups_per_cu, 8, idle_exit_us)
        h = C.c_void_p()
        check(ctx.lib.meso_serve_start(ctx.h, byref(cfg), C.byref(h)))
        self.h = h

    def submit(self, scene, cam, flags, sh):
        t = C.c_uint64()
        st = self.ctx.lib.meso_serve_submit(self.h, scene.h, byref(cam), flags,
                                            byref(sh) if sh else None, C.byref(t))
        if st == L.ERR_BUSY:
            return None
        check(st)
        return t.value

    def poll(self, ticket):
        d = C.c_int32()
        check(self.ctx.lib.meso_serve_poll(self.h, ticket, C.byref(d)))
        return d.value

    def wait(self, ticket, timeout_us=-1):
        lat = C.c_double()
        check(self.ctx.lib.meso_serve_wait(self.h, ticket, timeout_us, C.byref(lat)))
        return lat.value

    def take(self, ticket, width, height, index_bytes=4, parts=("index", "face", "t", "shade")):
        """Copy the named result arrays to host memory and release the
        request's HBM budget; arrays not named are returned as None."""
        n = width * height
        idx = np.empty(n, np.int64 if index_bytes == 8 else np.int32) if "index" in parts else None
        face = np.empty(n, np.uint8) if "face" in parts else None
        t = np.empty(n, np.float32) if "t" in parts else None
        shade = np.empty(n, np.float32) if "shade" in parts else None
        ptr = lambda a: a.ctypes.data if a is not None else None  # noqa: E731
        o = L.HostOutputs(C.sizeof(L.HostOutputs), index_bytes, ptr(idx), ptr(face), ptr(t),
                          ptr(shade))
        check(self.ctx.lib.meso_serve_take(self.h, ticket, byref(o)))
        return (idx.astype(np.int64) if idx is not None else None), face, t, shade

    def stats(self):
        s = L.ServeStats()
        s.size = C.sizeof(s)
        check(self.ctx.lib.meso_serve_stats_get(self.h, byref(s)))
        return {k: getattr(s, k) for k, _ in L.ServeStats._fields_ if k != "size"}

    def stop(self):
        if self.h:
            check(self.ctx.lib.meso_serve_stop(self.h))
            self.h = None


def edit_chunks(scene, xyz_buf, n_chunks, seed, stream=None):
    """Seeded sphere-toggle edits of the 32^3 chunks listed (int32 xyz triples
    on the device)."""
    check(scene.ctx.lib.meso_scene_edit_chunks(scene.h, xyz_buf.ptr, n_chunks, seed, stream))


def mesh_scene_chunks(ctx, scene, xyz_buf, n_chunks, mode, mt, stream=None):
    """Batched meshing of chunks read straight from a resident scene."""
    check(ctx.lib.meso_mesh_scene_chunks(ctx.h, scene.h, xyz_buf.ptr, n_chunks, mode,
                                         mt.counts.ptr, mt.prefix.ptr, mt.quads.ptr, mt.capacity,
                                         mt.total.ptr, mt.work.ptr, stream))


class ChunkPool:
    """The chunks around a camera, ranked by importance and kept meshed in
    fixed slots (meso_chunk_pool_*)."""

    def __init__(self, ctx, scene, capacity=16384, quads_per_slot=8192, forward_radius=24,
                 backward_radius=6, view_angle=120.0, max_loads=256, mode=L.MESH_GREEDY,
                 bake_directions=0):
        self.ctx, self.scene = ctx, scene
        self.desc = L.PoolDesc(C.sizeof(L.PoolDesc), capacity, quads_per_slot, forward_radius,
                               backward_radius, view_angle, max_loads, mode, bake_directions)
        h = C.c_void_p()
        check(ctx.lib.meso_chunk_pool_create(ctx.h, scene.h, byref(self.desc), C.byref(h)))
        self.h = h
        a = L.PoolArrays()
        a.size = C.sizeof(a)
        check(ctx.lib.meso_chunk_pool_arrays(self.h, byref(a)))
        self.arrays = a
        self.nc = a.chunks_per_axis

    def update(self, cam_chunk, forward, stream=None):
        cam = (C.c_int32 * 3)(*[int(v) for v in cam_chunk])
        fwd = (C.c_float * 3)(*[float(v) for v in forward])
        check(self.ctx.lib.meso_chunk_pool_update(self.h, C.cast(cam, C.c_void_p),
                                                  C.cast(fwd, C.c_void_p), stream))

    def remesh(self, xyz_buf, n_chunks, stream=None):
        check(self.ctx.lib.meso_chunk_pool_remesh(self.h, xyz_buf.ptr, n_chunks, stream))

    def stats(self, stream=None):
        s = L.PoolStats()
        s.size = C.sizeof(s)
        check(self.ctx.lib.meso_chunk_pool_stats(self.h, byref(s), stream))
        return {k: getattr(s, k) for k, _ in L.PoolStats._fields_ if k not in ("size", "pad")}

    def _d2h(self, ptr, dtype, count):
        out = np.empty(count, dtype)
        if count:
            check(self.ctx.lib.meso_memcpy_d2h(self.ctx.h, out.ctypes.data, ptr, out.nbytes, None))
        return out

    def state(self):
        """The pool's arrays, downloaded (waits for the context stream)."""
        cap, q = self.desc.capacity, self.desc.quads_per_slot
        a = self.arrays
        return {"lookup": self._d2h(a.lookup, np.uint32, self.nc ** 3),
                "slot_chunk": self._d2h(a.slot_chunk, np.uint32, cap),
                "slot_stamp": self._d2h(a.slot_stamp, np.uint32, cap),
                "slot_count": self._d2h(a.slot_count, np.uint32, cap),
                "quads": self._d2h(a.quads, np.uint32, cap * q),
                "load_xyz": self._d2h(a.load_xyz, np.int32, 3 * self.desc.max_loads),
                "load_slot": self._d2h(a.load_slot, np.uint32, self.desc.max_loads)}

    def destroy(self):
        if self.h:
            self.ctx.lib.meso_chunk_pool_destroy(self.h)
            self.h = None


class Raster:
    """Quad visibility of a chunk pool's resident quads (meso_raster_*)."""

    def __init__(self, ctx, max_width, max_height):
        self.ctx = ctx
        h = C.c_void_p()
        check(ctx.lib.meso_raster_create(ctx.h, max_width, max_height, C.byref(h)))
        self.h = h

    def draw(self, pool, cam, sh, targets, stream=None):
        check(self.ctx.lib.meso_raster_pool(self.h, pool.h, byref(cam), byref(sh) if sh else None,
                                            byref(targets.out), stream))

    def stats(self, stream=None):
        s = L.RasterStats()
        s.size = C.sizeof(s)
        check(self.ctx.lib.meso_raster_stats_get(self.h, byref(s), stream))
        return {k: getattr(s, k) for k, _ in L.RasterStats._fields_ if k != "size"}

    def destroy(self):
        if self.h:
            self.ctx.lib.meso_raster_destroy(self.h)
            self.h = None
