"""Chunk streaming around a moving camera (meso_chunk_pool_*).

Two measurements:
  fly-through: the 2048^3 two-level cave volume; the camera circles the
    middle of the volume about one chunk per update, and each update ranks
    the chunks of
    its view sphere (radius 24 chunks forward, 6 backward, 120 degree cone,
    the view snapped to the nearest of 256 baked directions),
    loads at most 256 of them (the reference's default load limit) into a
    16,384-slot pool (the reference's default MaxChunkCount), evicting the
    least important, and greedy-meshes them into their slots.  The untimed
    warm-up updates fill the pool, so the timed ones also evict.  Timed per
    update with events; reported as ms per update and Mvoxels meshed per
    second.  A second pool loads all 16,384 slots in one update (a cold
    start).
  sample: the 512^3 heightfield with a 4,096-slot pool and radius 16, two
    updates; the same updates run on the CPU oracle in bench.py's CPU leg.
  raster: the fly-through pool's resident quads drawn at 1920x1080
    (meso_raster_pool), timed beside the raymarcher's primary rays on the
    same camera; outside the resident radius the raymarcher sees geometry
    the pool does not hold, so agreement counts every hit pixel.
"""
import time

from . import _lib as L
from . import runtime as R
from . import workloads as W

FLY = dict(capacity=16384, quads_per_slot=8192, forward_radius=24, backward_radius=6,
           view_angle=120.0, max_loads=256, bake_directions=256)
SAMPLE = dict(capacity=4096, quads_per_slot=8192, forward_radius=16, backward_radius=6,
              view_angle=120.0, max_loads=256)
SAMPLE_PATH = [((8, 4, 8), (1.0, -0.2, 0.5)), ((9, 4, 8), (1.0, -0.2, 0.6))]


def fly_path(k):
    """camera chunk and forward vector of update k of the fly-through: a
    circle of radius 20 chunks around the middle of the 64^3-chunk volume,
    about one chunk per update, looking along the circle"""
    import math
    th = k / 20.0
    cam = (32 + int(round(20 * math.cos(th))), 32, 32 + int(round(20 * math.sin(th))))
    return cam, (-math.sin(th), 0.15, math.cos(th))


def _timed_update(ctx, pool, cam, fwd):
    e0, e1 = R.Event(ctx), R.Event(ctx)
    e0.record()
    pool.update(cam, fwd)
    e1.record()
    e1.sync()
    return e0.elapsed_ms(e1)


def run(ctx, updates=24, warmup=72, log=print):
    cfg = W.CAVES
    t0 = time.time()
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, cfg["n"], cfg["seed"])
    ctx.sync()
    log("streaming: 2048^3 caves generated in %.2fs" % (time.time() - t0))
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **FLY)
    for k in range(warmup):
        pool.update(*fly_path(k))
    ctx.sync()
    ms, loads, evicted, cands = [], 0, 0, []
    for k in range(warmup, warmup + updates):
        ms.append(_timed_update(ctx, pool, *fly_path(k)))
        st = pool.stats()
        loads += st["n_loads"]
        evicted += st["n_evicted"]
        cands.append(st["n_candidates"])
    st = pool.stats()
    med = sorted(ms)[len(ms) // 2]
    total_ms = sum(ms)
    fly = dict(updates=updates, ms_per_update_median=round(med, 4),
               ms_per_update_mean=round(total_ms / updates, 4),
               loads_per_update=round(loads / updates, 1),
               evicted_per_update=round(evicted / updates, 1),
               candidates_per_update=round(sum(cands) / updates, 1),
               mvox_per_s=round(loads * 32768 / (total_ms * 1e-3) / 1e6, 1),
               resident=st["n_resident"], resident_quads=st["n_quads"],
               truncated=st["n_truncated"], pool_mib=round(st["device_bytes"] / 2 ** 20, 1),
               params=FLY)
    log("streaming: fly-through %.3f ms/update (median), %.0f loads, %.0f candidates, "
        "%.0f Mvox/s" % (med, fly["loads_per_update"], fly["candidates_per_update"],
                         fly["mvox_per_s"]))
    raster = draw_leg(ctx, sc, pool, *fly_path(warmup + updates - 1), log=log)
    pool.destroy()
    # cold start: every slot in one update
    cold = dict(FLY, max_loads=FLY["capacity"])
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **cold)
    pool.update(*fly_path(0))  # untimed first use
    pool.destroy()
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **cold)
    ms0 = _timed_update(ctx, pool, *fly_path(0))
    st = pool.stats()
    fill = dict(ms=round(ms0, 3), loads=st["n_loads"], candidates=st["n_candidates"],
                mvox_per_s=round(st["n_loads"] * 32768 / (ms0 * 1e-3) / 1e6, 1))
    log("streaming: cold start %d loads in %.3f ms (%.0f Mvox/s)" % (
        fill["loads"], ms0, fill["mvox_per_s"]))
    pool.destroy()
    sc.destroy()
    # the CPU-comparable sample
    t = W.TERRAIN
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, t["n"], t["seed"])
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **SAMPLE)
    pool.update(*SAMPLE_PATH[0])  # untimed first use
    pool.destroy()
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **SAMPLE)
    sample_ms, sample_loads = 0.0, 0
    for cam, fwd in SAMPLE_PATH:
        sample_ms += _timed_update(ctx, pool, cam, fwd)
        sample_loads += pool.stats()["n_loads"]
    pool.destroy()
    sc.destroy()
    sample = dict(ms=round(sample_ms, 3), loads=sample_loads,
                  mvox_per_s=round(sample_loads * 32768 / (sample_ms * 1e-3) / 1e6, 1),
                  workload="512^3 heightfield brick map, %d slots, radius %d, %d updates" % (
                      SAMPLE["capacity"], SAMPLE["forward_radius"], len(SAMPLE_PATH)))
    return dict(fly_through=fly, cold_start=fill, sample=sample, raster=raster)


def draw_leg(ctx, sc, pool, cam_chunk, fwd, width=1920, height=1080, frames=10, log=print):
    """The pool's resident quads drawn at 1920x1080 from the camera of the
    last fly-through update (meso_raster_pool), against the raymarcher's
    primary rays (flags 0) on the same camera and scene."""
    eye = [c * 32 + 16.0 for c in cam_chunk]
    target = [eye[i] + fwd[i] for i in range(3)]
    cam = R.camera(eye, target, [0.0, 1.0, 0.0], 70.0, width, height)
    sh = R.shading(W.LIGHT)
    rast = R.Raster(ctx, width, height)
    tg = R.RayTargets(ctx, width, height, index_bytes=8)
    rast.draw(pool, cam, sh, tg)  # untimed first use
    ctx.sync()

    def timed(fn):
        e0, e1 = R.Event(ctx), R.Event(ctx)
        e0.record()
        for _ in range(frames):
            fn()
        e1.record()
        e1.sync()
        return e0.elapsed_ms(e1) / frames

    ms = timed(lambda: rast.draw(pool, cam, sh, tg))
    st = rast.stats()
    gi, gf, _, _ = tg.results()
    tr = R.RayTargets(ctx, width, height, index_bytes=8)
    R.raymarch(ctx, sc, cam, 0, sh, tr)
    ctx.sync()
    ms_rm = timed(lambda: R.raymarch(ctx, sc, cam, 0, sh, tr))
    di, df, _, _ = tr.results()
    hit = (gi >= 0) | (di >= 0)
    agree = float(((gi == di) & (gf == df))[hit].mean()) if hit.any() else 1.0
    rast.destroy()
    out = dict(ms=round(ms, 4), raymarch_primary_ms=round(ms_rm, 4),
               resident_quads=pool.stats()["n_quads"], hit_px=int((gi >= 0).sum()),
               agree_with_raymarch=round(agree, 5), **st)
    log("streaming: raster %.3f ms at %dx%d (%d small, %d large quads, %d segments); "
        "raymarch primary %.3f ms; agree %.4f" % (ms, width, height, st["small_quads"],
                                                   st["large_quads"], st["segments"], ms_rm,
                                                   agree))
    return out
