"""Every raymarch launch shape gives the oracle's bits and step counts.

The five tiled shapes, the persistent waves on 8x8, 16x4 and 32x2 packets
(variants 5-7) and the waves over pairs of packets whose secondary rays are
shared over the wave (variants 8-9), at the device's full occupancy and at a
small grid that makes every wave serve several bands, are forced through the
autotuner's cache file, then run through meso_raymarch twice: with the step
counters on (the counting instantiation, whose counts must equal the
oracle's too) and with them off, the instantiation that timed frames run.
"""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu

# (variant, tile_w, tile_h, grid)
SHAPES = [(0, 16, 16, 0), (1, 32, 8, 0), (2, 8, 8, 0), (3, 16, 8, 0), (4, 32, 16, 0),
          (5, 8, 8, 0), (5, 8, 8, 64), (6, 16, 4, 0), (7, 32, 2, 64), (8, 8, 8, 0),
          (8, 8, 8, 64), (9, 16, 4, 64)]


def _force(ctx, tmp_path, kind, n, w, h, flags, shape):
    v, tw, th, grid = shape
    key = "ray:%d:%d:%dx%d:%d" % (kind, n, w, h, flags)
    p = tmp_path / "tune.txt"
    p.write_text("%s %d %d %d %d %d 1 1\n" % (key, v, 64 if v >= 5 else tw * th, grid, tw, th))
    L.check(ctx.lib.meso_autotune_load(ctx.h, str(p).encode()))


def _check(ctx, tg, ref, what, count=True):
    idx, face, t, shade = tg.results()
    assert np.array_equal(idx, ref[0]), what
    assert np.array_equal(face, ref[1]), what
    assert np.array_equal(t.view(np.uint32), ref[2].view(np.uint32)), what
    assert np.array_equal(shade.view(np.uint32), ref[3].view(np.uint32)), what
    if not count:
        return
    cnt = tg.counts()
    rc = ref[4].as_dict()
    for k in rc:
        assert cnt[k] == rc[k], (what, k, cnt[k], rc[k])


def _fresh(tg):
    """poison the outputs so a launch that skips a pixel cannot pass"""
    for b in (tg.index, tg.face, tg.t, tg.shade):
        b.fill(0xA5)


@pytest.mark.parametrize("count", [True, False], ids=["count", "timed"])
@pytest.mark.parametrize("shape", SHAPES)
def test_brickmap_shapes_match_oracle(ctx, tmp_path, shape, count):
    n, w, h = 256, 328, 200  # frame size used by no other test: its cache key is this test's
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, n, 2)
    cp = W.terrain_camera(n, w, h)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    _force(ctx, tmp_path, L.SCENE_BRICKMAP_U16, n, w, h, L.RAY_SHADOW, shape)
    tg = R.RayTargets(ctx, w, h, counters=True)
    _fresh(tg)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW | (L.RAY_COUNT if count else 0), R.shading(W.LIGHT), tg)
    ctx.sync()
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, n, 2)
    osc = O.brickmap_scene(n, slots, masks, vox)
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading(W.LIGHT), threads=8)
    _check(ctx, tg, ref, "variant %s" % (shape,), count)
    sc.destroy()


@pytest.mark.parametrize("count", [True, False], ids=["count", "timed"])
@pytest.mark.parametrize("shape", [(0, 16, 16, 0), (2, 8, 8, 0), (5, 8, 8, 0), (5, 8, 8, 24),
                                   (6, 16, 4, 24), (7, 32, 2, 0), (8, 8, 8, 0), (8, 8, 8, 24),
                                   (9, 16, 4, 0)])
def test_persistent_two_level_ao_and_dense(ctx, tmp_path, shape, count):
    n, w, h = 256, 136, 72
    cflag = L.RAY_COUNT if count else 0
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, 5)
    cp = W.caves_camera(n, w, h)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    flags = L.RAY_SHADOW | L.RAY_AO
    _force(ctx, tmp_path, L.SCENE_TWO_LEVEL_U16, n, w, h, flags, shape)
    tg = R.RayTargets(ctx, w, h, index_bytes=8, counters=True)
    _fresh(tg)
    R.raymarch(ctx, sc, cam, flags | cflag, R.shading(W.LIGHT, 9, 4, 16.0), tg)
    ctx.sync()
    slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, 5)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    osc = O.two_level_scene(n, top, ns, m2, v2)
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    ref = O.raymarch(osc, ocam, flags, O.shading(W.LIGHT, 9, 4, 16.0), threads=8)
    _check(ctx, tg, ref, "two-level %s" % (shape,), count)
    sc.destroy()

    # the dense sphere grid, an odd frame size (partial packets at both edges)
    w, h = 203, 117
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    cp = W.spheres_camera(width=w, height=h)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    _force(ctx, tmp_path, L.SCENE_DENSE_U8, 128, w, h, L.RAY_SHADOW, shape)
    tg = R.RayTargets(ctx, w, h, counters=True)
    _fresh(tg)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW | cflag, R.shading(W.LIGHT), tg)
    ctx.sync()
    osc = O.dense_scene(O.spheres_dense(1, 128, 8).ravel())
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading(W.LIGHT), threads=8)
    _check(ctx, tg, ref, "dense %s" % (shape,), count)
    sc.destroy()


def test_pin_validates_shapes(ctx):
    import ctypes as C
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    cam = R.camera([300.0, 250.0, 200.0], [64.0, 64.0, 64.0], [0.0, 0.0, 1.0], 60.0, 64, 48)
    t = L.TuneResult()
    t.size = C.sizeof(t)
    t.variant, t.tile_w, t.tile_h = 2, 16, 16  # variant 2 has 8x8 tiles
    assert ctx.lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), 0, L.byref(t)) == L.ERR_ARG
    t.variant, t.grid = 5, 33 * ctx.device_info().compute_units
    assert ctx.lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), 0, L.byref(t)) == L.ERR_ARG
    t.variant = 10
    assert ctx.lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), 0, L.byref(t)) == L.ERR_ARG
    # a valid pin renders the same bits as the default shape
    tg0 = R.RayTargets(ctx, 64, 48)
    R.raymarch(ctx, sc, cam, 0, None, tg0)
    t.variant, t.grid, t.tile_w, t.tile_h = 5, 8, 8, 8
    L.check(ctx.lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), 0, L.byref(t)))
    tg1 = R.RayTargets(ctx, 64, 48)
    R.raymarch(ctx, sc, cam, 0, None, tg1)
    ctx.sync()
    for a, b in zip(tg0.results(), tg1.results()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    sc.destroy()
