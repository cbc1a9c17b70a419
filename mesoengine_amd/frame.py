"""Frame loop over the 2048^3 two-level cave volume.

Each frame:
  1. edit: 1% of the 32^3 chunks (2,621 of 262,144, a seeded set per frame)
     receive seeded sphere-toggle edits in place (meso_scene_edit_chunks);
  2. remesh: the dirty chunks are greedy-meshed straight from the resident
     scene in one persistent launch (meso_mesh_scene_chunks);
  3. render: one 3840x2160 frame of primary rays + a shadow ray per lit hit +
     4 cosine-hemisphere AO rays per hit (meso_raymarch with SHADOW | AO).
headline() times K frames back to back with each frame's remesh on a second
stream beside its render (the remesh and the render only read the edited
scene; the next frame's edit waits for both), the render alone (the
roofline's launch duration), and the three steps one after another on one
stream (the split).  under_serving_load() runs the frames while a serving
session (the 16-scene request stream of the serving workload) runs on the
server's stream, and reports the frame split and the served latency.
"""
import threading
import time

import numpy as np

from . import _lib as L
from . import runtime as R
from . import serving
from . import workloads as W


def dirty_chunks(n, frame, seed, frac=0.01):
    """Seeded set of unique dirty chunks for one frame, int32 (x, y, z) rows."""
    nc = -(-n // 32)
    total = nc ** 3
    count = int(total * frac)
    rng = np.random.default_rng([seed, frame])
    flat = rng.choice(total, size=count, replace=False)
    return np.stack([flat % nc, (flat // nc) % nc, flat // (nc * nc)], 1).astype(np.int32)


def frame_edit(n, frame, seed):
    """The dirty chunks and edit seed of frame `frame`.  Frames come in
    pairs that apply the same edit: an edit toggles occupancy inside its
    spheres (solid -> empty, empty -> solid), so the second frame of a pair
    restores the first's occupancy and the scene does not drift over a run;
    every frame still edits and remeshes 1% of the chunks."""
    k = frame // 2
    return dirty_chunks(n, k, seed), seed * 7919 + k


def frame_bytes(cnt, index_bytes=8):
    """Algorithmic bytes of one rendered frame (the path's per-ray byte
    model): 4 B per top-level and node pointer read, 64 B per non-empty brick
    entered (its occupancy mask), 2 B material per hit, and per pixel the
    outputs (index, 1 B face, 4 B t, 4 B shade)."""
    return (4 * cnt["n_top"] + 4 * cnt["n_node"] + 64 * cnt["n_brickmask"] + 2 * cnt["hits"]
            + (index_bytes + 9) * cnt["rays"])


class _Frames:
    def __init__(self, ctx, cfg, seed, log):
        self.ctx = ctx
        self.cfg = cfg
        n = cfg["n"]
        t = time.time()
        self.scene = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, cfg["seed"])
        ctx.sync()
        i = self.scene.info()
        self.gen_s = time.time() - t
        self.info = dict(n=n, bricks=int(i.n_bricks), nodes=int(i.n_nodes),
                         device_mb=round(i.device_bytes / 1e6, 1),
                         fill=round(i.solid_voxels / n ** 3, 5), generate_s=round(self.gen_s, 3))
        log("frame scene %d^3 caves: %d bricks in %d nodes, %.0f MB, %.4f filled, built in %.2fs" % (
            n, i.n_bricks, i.n_nodes, i.device_bytes / 1e6, i.solid_voxels / n ** 3, self.gen_s))
        cp = W.caves_camera(n, cfg["width"], cfg["height"])
        self.cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], cfg["width"],
                            cfg["height"])
        self.sh = R.shading(W.LIGHT, seed, cfg["ao_samples"], cfg["ao_radius"])
        self.flags = L.RAY_SHADOW | L.RAY_AO
        self.tg = R.RayTargets(ctx, cfg["width"], cfg["height"], index_bytes=8, counters=True)
        self.seed = seed
        self.n_dirty = len(dirty_chunks(n, 0, seed))
        self.mt = R.MeshTargets(ctx, self.n_dirty, 24 << 20)
        self.ev = [R.Event(ctx) for _ in range(4)]

    def tune(self):
        """autotune the frame's launch shape (kept by the library for this key)"""
        import ctypes as C
        t = L.TuneResult()
        t.size = C.sizeof(t)
        L.check(self.ctx.lib.meso_autotune_raymarch(self.ctx.h, self.scene.h, L.byref(self.cam),
                                                    self.flags, L.byref(self.sh),
                                                    L.byref(self.tg.out), L.byref(t)))
        return dict(variant=t.variant, grid=t.grid, tile=[t.tile_w, t.tile_h],
                    best_ms=round(t.best_ms, 4), default_ms=round(t.default_ms, 4))

    def pin(self, variant, waves_per_cu=0):
        """pin the render's launch shape: a tiled variant (0..4), or persistent
        waves (5) on `waves_per_cu` of the 32 wave slots of each CU"""
        import ctypes as C
        t = L.TuneResult()
        t.size = C.sizeof(t)
        tiles = [(16, 16), (32, 8), (8, 8), (16, 8), (32, 16), (8, 8)]
        t.variant = variant
        t.tile_w, t.tile_h = tiles[variant]
        t.grid = waves_per_cu * self.ctx.device_info().compute_units if variant == 5 else 0
        L.check(self.ctx.lib.meso_autotune_pin_raymarch(self.ctx.h, self.scene.h,
                                                        L.byref(self.cam), self.flags,
                                                        L.byref(t)))

    def count(self):
        """step counts of one render of the current scene"""
        self.tg.counters.zero()
        R.raymarch(self.ctx, self.scene, self.cam, self.flags | L.RAY_COUNT, self.sh, self.tg)
        self.ctx.sync()
        return self.tg.counts()

    def _edits(self, frames, first):
        plan = [frame_edit(self.cfg["n"], first + f, self.seed) for f in range(frames)]
        lists = [self.ctx.upload(xyz.ravel()) for xyz, _ in plan]
        self.ctx.sync()
        return lists, [s for _, s in plan]

    def run(self, frames, first=0):
        """Time `frames` frames; returns per-frame (edit, remesh, render) ms and quads."""
        lists, seeds = self._edits(frames, first)
        out = []
        e = self.ev
        for f in range(frames):
            e[0].record()
            R.edit_chunks(self.scene, lists[f], self.n_dirty, seeds[f])
            e[1].record()
            R.mesh_scene_chunks(self.ctx, self.scene, lists[f], self.n_dirty, L.MESH_GREEDY,
                                self.mt)
            e[2].record()
            R.raymarch(self.ctx, self.scene, self.cam, self.flags, self.sh, self.tg)
            e[3].record()
            e[3].sync()
            total = int(self.mt.total.download(np.uint64, 1)[0])
            out.append((e[0].elapsed_ms(e[1]), e[1].elapsed_ms(e[2]), e[2].elapsed_ms(e[3]),
                        total))
        return out

    def run_overlapped(self, frames, stream2, first=0):
        """`frames` frames back to back, timed as one interval.  Each frame's
        edit runs on the context stream; its remesh runs on `stream2` beside
        its render (both only read the edited scene); the next frame's edit
        waits for both.  Returns ms per frame."""
        lists, seeds = self._edits(frames, first)
        L.check(self.ctx.lib.meso_stream_sync(self.ctx.h, stream2))
        a, b, ev_edit, ev_mesh = (R.Event(self.ctx) for _ in range(4))
        a.record()
        for f in range(frames):
            R.edit_chunks(self.scene, lists[f], self.n_dirty, seeds[f])
            ev_edit.record()
            L.check(self.ctx.lib.meso_stream_wait_event(self.ctx.h, stream2, ev_edit.h))
            R.mesh_scene_chunks(self.ctx, self.scene, lists[f], self.n_dirty, L.MESH_GREEDY,
                                self.mt, stream2)
            ev_mesh.record(stream2)
            R.raymarch(self.ctx, self.scene, self.cam, self.flags, self.sh, self.tg)
            L.check(self.ctx.lib.meso_stream_wait_event(self.ctx.h, None, ev_mesh.h))
        b.record()
        b.sync()
        return a.elapsed_ms(b) / frames

    def count_frames(self, frames, first=0):
        """The step counts summed over `frames` frames (edits applied as in
        the timed frames, renders with the counters on, untimed): replayed
        from the same occupancy, they are the counts of the timed frames."""
        lists, seeds = self._edits(frames, first)
        self.tg.counters.zero()
        for f in range(frames):
            R.edit_chunks(self.scene, lists[f], self.n_dirty, seeds[f])
            R.raymarch(self.ctx, self.scene, self.cam, self.flags | L.RAY_COUNT, self.sh, self.tg)
        self.ctx.sync()
        return self.tg.counts()

    def render_alone(self, k):
        """k renders back to back on the context stream (the stream the
        kernel runs on): the render kernel's average launch duration"""
        for _ in range(2):
            R.raymarch(self.ctx, self.scene, self.cam, self.flags, self.sh, self.tg)
        e0, e1 = R.Event(self.ctx), R.Event(self.ctx)
        e0.record()
        for _ in range(k):
            R.raymarch(self.ctx, self.scene, self.cam, self.flags, self.sh, self.tg)
        e1.record()
        e1.sync()
        return e0.elapsed_ms(e1) / k

    def close(self):
        self.scene.destroy()


def _summary(rows):
    a = np.array([r[:3] for r in rows], np.float64)
    med = np.median(a, axis=0)
    return dict(frames=len(rows), edit_ms=round(float(med[0]), 4),
                remesh_ms=round(float(med[1]), 4), render_ms=round(float(med[2]), 4),
                frame_ms=round(float(np.median(a.sum(1))), 4),
                frame_ms_max=round(float(a.sum(1).max()), 4),
                quads_per_frame=int(np.median([r[3] for r in rows])))


def _under_load(ctx, fr, scenes, reqs, budget_bytes, frames, rays, log, label):
    """frames rendered while the serving stream runs; frame split plus the
    served latency"""
    sh = R.shading(W.LIGHT)
    result = {}

    def serve():
        try:
            # the server's persistent kernel on 8 wave slots per CU (2
            # workgroups of 4 waves), so the render keeps the rest
            result["lat"], result["makespan"], result["stats"] = serving.served_run(
                ctx, scenes, reqs, sh, budget_bytes, True, workgroups_per_cu=2)
        except Exception as ex:  # surfaced after join
            result["error"] = ex

    th = threading.Thread(target=serve)
    th.start()
    rows = []
    f = 0
    while th.is_alive() or len(rows) < frames:
        rows += fr.run(1, first=20_000 + f)
        f += 1
        if len(rows) >= 4 * frames and not th.is_alive():
            break
    th.join()
    if "error" in result:
        raise result["error"]
    out = _summary(rows)
    out["render_mrays"] = round(rays / (out["render_ms"] * 1e-3) / 1e6, 1)
    lat = result["lat"]
    out["serving"] = dict(p50_ms=round(float(np.percentile(lat, 50) * 1e3), 3),
                          p99_ms=round(float(np.percentile(lat, 99) * 1e3), 3),
                          launches=int(result["stats"]["launches"]))
    out["render_shape"] = label
    log("frame under serving load, %s: edit %.3f ms, remesh %.3f ms, render %.3f ms, frame "
        "%.3f ms (max %.3f); served p50 %.2f p99 %.2f ms" % (
            label, out["edit_ms"], out["remesh_ms"], out["render_ms"], out["frame_ms"],
            out["frame_ms_max"], out["serving"]["p50_ms"], out["serving"]["p99_ms"]))
    return out


def headline(ctx, steps=20, warmup=3, seed=5, log=print):
    """The 2048^3 caves frame loop as the benchmark's headline: K frames back
    to back (edit, then remesh beside the render), the render alone for the
    roofline, and the frame split.  Returns (result, frames object); the
    caller closes the frames object."""
    cfg = W.CAVES
    fr = _Frames(ctx, cfg, seed, log)
    try:
        tune = fr.tune()
        log("frame render tune: %s" % tune)
        import ctypes as C
        shape = L.TuneResult()
        shape.size = C.sizeof(shape)
        shape.variant, shape.grid = tune["variant"], tune["grid"]
        shape.tile_w, shape.tile_h = tune["tile"]
        if not R.matches_tiled_frame(ctx, fr.scene, fr.cam, fr.flags, fr.sh, fr.tg, shape):
            raise RuntimeError("the tuned render shape's frame differs from the 16x16-tile frame")
        log("frame render: the tuned shape's frame equals the 16x16-tile frame")
        cnt = fr.count()
        rays = cnt["rays"] + cnt["shadow_rays"] + cnt["ao_rays"]
        alg = frame_bytes(cnt)
        steps += steps % 2  # whole pairs: the timed frames end on the start occupancy
        warmup += warmup % 2
        s2 = C.c_void_p()
        L.check(ctx.lib.meso_stream_create(ctx.h, C.byref(s2)))
        try:
            fr.run_overlapped(warmup, s2, first=10_000)
            ms = fr.run_overlapped(steps, s2, first=0)
        finally:
            L.check(ctx.lib.meso_stream_destroy(ctx.h, s2))
        timed = fr.count_frames(steps, first=0)
        rays_timed = timed["rays"] + timed["shadow_rays"] + timed["ao_rays"]
        render_ms = fr.render_alone(steps)
        split = _summary(fr.run(min(steps, 8), first=1_000))
        remesh_bytes = 65536 * fr.n_dirty + 4 * split["quads_per_frame"] + 8 * fr.n_dirty
        out = dict(
            workload="2048^3 two-level caves (u16, 8^3 bricks under 16^3-brick nodes), 3840x2160 "
                     "primary + shadow + 4 AO rays per hit; 1% of 32^3 chunks edited and "
                     "greedy-remeshed per frame",
            scene=fr.info, dirty_chunks=fr.n_dirty, counts=cnt,
            rays_per_frame=int(round(rays_timed / steps)), rays_start_state=rays,
            render_tune=tune, frames=steps, rays_timed_frames=rays_timed,
            ms_per_frame=ms, mrays=rays_timed / steps / (ms * 1e-3) / 1e6,
            render_ms=render_ms, render_mrays=rays / (render_ms * 1e-3) / 1e6,
            render_alg_bytes=alg, render_alg_gbs=alg / (render_ms * 1e-3) / 1e9,
            split_one_stream=split,
            remesh_mvox_s=round(fr.n_dirty * 32768 / (split["remesh_ms"] * 1e-3) / 1e6, 1),
            remesh_alg_gbs=round(remesh_bytes / (split["remesh_ms"] * 1e-3) / 1e9, 1))
        log("frame loop: %.3f ms per frame (edit, then remesh beside the render), %.0f Mrays/s; "
            "render alone %.3f ms (%.0f GB/s algorithmic); one stream: edit %.3f, remesh %.3f, "
            "render %.3f ms" % (ms, out["mrays"], render_ms, out["render_alg_gbs"],
                                split["edit_ms"], split["remesh_ms"], split["render_ms"]))
        return out, fr
    except Exception:
        fr.close()
        raise


def under_serving_load(ctx, fr, rays, frames=8, serve_qps=1000.0, serve_requests=400,
                       budget_bytes=512 << 20, log=print):
    """The frame loop of `fr` while the serving stream (the 16-scene request
    stream of the serving workload) runs on the server's own stream, with
    three render shapes: the one tuned for the frame alone (persistent waves
    on every slot), 16x16 tiles (slots return to the dispatcher as tiles
    finish), persistent waves on half of the slots."""
    edges = W.serve_scenes(4)
    scenes = serving.make_scenes(ctx, edges, 4)
    reqs = W.serve_requests(edges, serve_requests, serve_qps, 4)
    loaded = {}
    try:
        for label, pin in (("tuned", None), ("tiled_16x16", (0, 0)),
                           ("persistent_16_of_32_slots", (5, 16))):
            if pin:
                fr.pin(*pin)
            loaded[label] = _under_load(ctx, fr, scenes, reqs, budget_bytes, frames, rays,
                                        log, label)
            loaded[label]["serving"]["qps"] = serve_qps
            loaded[label]["serving"]["requests"] = serve_requests
    finally:
        for s in scenes:
            s.destroy()
    return loaded
