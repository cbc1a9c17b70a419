"""Quad visibility on the GPU against the oracle's: the chunk pool's quads
drawn by meso_raster_pool give the oracle's index, face, t and shade bit for
bit (small quads tested by one lane, medium ones by a wave, large ones cut
into segments, quads clipped at the eye), and agree with the GPU
raymarcher's primary hits.  A close-up frame makes quads of every size."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = [(L.SCENE_BRICKMAP_U16, L.GEN_HEIGHTFIELD, 128, 2, (-30.0, 90.0, -20.0), (64.0, 30.0, 64.0),
          160, 120),
         (L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, 256, 5, (128.0, 140.0, 120.0), (200.0, 100.0, 180.0),
          320, 200),
         (L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, 256, 5, (-40.0, 300.0, -40.0), (128.0, 100.0, 128.0),
          203, 117),
         # a floor slab (y < 8) with a block on it, seen from just above: 32x32
         # greedy quads that cover thousands of pixels
         (L.SCENE_BRICKMAP_U16, "slab", 64, 0, (20.0, 9.5, 6.0), (40.0, 4.0, 50.0), 400, 300)]


def _slab(ctx, n):
    g = np.zeros((n, n, n), np.uint16)  # (z, y, x)
    g[:, :8, :] = 1
    g[20:28, 8:14, 30:40] = 2
    slots, masks, vox = O.brickmap_from_dense(g)
    return ctx.scene_brickmap(n, slots, masks, vox)


@pytest.mark.parametrize("kind,gen,n,seed,eye,target,w,h", CASES)
def test_raster_matches_oracle_and_raymarch(ctx, kind, gen, n, seed, eye, target, w, h):
    sc = _slab(ctx, n) if gen == "slab" else ctx.scene_generate(gen, kind, n, seed)
    nc = -(-n // 32)
    R_ = int(np.ceil(np.sqrt(3) * nc)) + 1
    per = 65536
    pool = R.ChunkPool(ctx, sc, capacity=nc ** 3, quads_per_slot=per, forward_radius=R_,
                       backward_radius=R_, view_angle=120.0, max_loads=nc ** 3)
    pool.update((nc // 2, nc // 2, nc // 2), (1.0, 0.0, 0.0))
    st = pool.state()
    assert pool.stats()["n_truncated"] == 0
    cam = R.camera(list(eye), list(target), [0.0, 1.0, 0.0], 70.0, w, h)
    rast = R.Raster(ctx, 400, 300)
    tg = R.RayTargets(ctx, w, h)
    sh = R.shading(W.LIGHT)
    rast.draw(pool, cam, sh, tg)
    ctx.sync()
    gi, gf, gt, gs = tg.results()
    rs = rast.stats()
    assert rs["small_quads"] + rs["large_quads"] > 0 and rs["overflow_quads"] == 0
    if gen == "slab":
        assert rs["segments"] > 0, rs  # quads of more than 4096 pixels
    else:
        assert rs["small_quads"] > 0 and rs["large_quads"] > 0, rs
    ocam = O.camera(list(eye), list(target), [0.0, 1.0, 0.0], 70.0, w, h)
    oi, of, ot, os_, _ = O.raster_pool(ocam, n, nc, st["slot_chunk"], st["slot_count"],
                                       st["quads"], per, O.shading(W.LIGHT).light)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gf, of)
    assert np.array_equal(gt.view(np.uint32), ot.view(np.uint32))
    assert np.array_equal(gs.view(np.uint32), os_.view(np.uint32))
    # the raymarcher's primary hits on the same frame
    tr = R.RayTargets(ctx, w, h)
    R.raymarch(ctx, sc, cam, 0, sh, tr)
    ctx.sync()
    di, df, dt, ds = tr.results()
    same = (di == gi) & (df == gf)
    assert (di >= 0).mean() > 0.2 and same.mean() >= 0.995, same.mean()
    assert np.array_equal(dt.view(np.uint32)[same], gt.view(np.uint32)[same])
    rast.destroy()
    pool.destroy()
    sc.destroy()


def test_raster_argument_checks(ctx):
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, 128, 2)
    pool = R.ChunkPool(ctx, sc, capacity=64, quads_per_slot=1024, forward_radius=4,
                       backward_radius=4, view_angle=120.0, max_loads=64)
    rast = R.Raster(ctx, 64, 64)
    h = C.c_void_p()
    assert ctx.lib.meso_raster_create(ctx.h, 0, 64, C.byref(h)) == L.ERR_ARG
    assert ctx.lib.meso_raster_create(ctx.h, 64, 20000, C.byref(h)) == L.ERR_ARG
    tg = R.RayTargets(ctx, 80, 60)
    cam = R.camera([-30.0, 90.0, -20.0], [64.0, 30.0, 64.0], [0.0, 1.0, 0.0], 70.0, 80, 60)
    st = ctx.lib.meso_raster_pool(rast.h, pool.h, L.byref(cam), None, L.byref(tg.out), None)
    assert st == L.ERR_ARG  # larger than the raster's 64 x 64
    cam = R.camera([-30.0, 90.0, -20.0], [64.0, 30.0, 64.0], [0.0, 1.0, 0.0], 70.0, 64, 48)
    bad = L.RayOutputs()
    bad.size = C.sizeof(bad)
    bad.index_bytes = 4
    assert ctx.lib.meso_raster_pool(rast.h, pool.h, L.byref(cam), None, L.byref(bad), None) \
        == L.ERR_ARG
    # an empty pool draws nothing
    tg = R.RayTargets(ctx, 64, 48)
    rast.draw(pool, cam, None, tg)
    ctx.sync()
    idx, face, t, shade = tg.results()
    assert (idx == -1).all() and (face == 0xFF).all() and np.isinf(t).all() and (shade == 0).all()
    rast.destroy()
    pool.destroy()
    sc.destroy()
