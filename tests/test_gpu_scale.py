"""The timed launch shapes and the large workloads, against the oracle.

- The 1920x1080 heightfield frame (primary + shadow) in the launch shapes
  the benchmark pins, with the step counters off (the instantiation timed
  frames run), one frame on the library stream and one on a second stream.
- The 2048^3 two-level cave volume: a sample of its bricks against the
  oracle's per-brick generator, and bands of the 3840x2160 primary + shadow
  + 4 AO frame, rendered in the autotuned shape, against the oracle's DDA on
  the downloaded volume.  Hit indices there exceed 2^31, where the
  reference's int32 FVoxelMathHelper::Convert3DTo1D
  (Runtimes/Helper/VoxelMathHelper.h:73-81) would overflow.
- One served request on each of three resident scenes with edges 256, 520
  and 1024, against the oracle on the downloaded scene.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def _pin(ctx, sc, cam, flags, variant, grid, tile=(8, 8)):
    t = L.TuneResult()
    t.size = C.sizeof(t)
    t.variant, t.grid = variant, grid
    t.tile_w, t.tile_h = tile
    L.check(ctx.lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), flags, L.byref(t)))


def _poison(tg):
    for b in (tg.index, tg.face, tg.t, tg.shade):
        b.fill(0xA5)


def _equal(got, ref, what, rows=None):
    """bit equality of (index, face, t, shade), optionally over pixel rows"""
    sl = slice(None) if rows is None else slice(rows[0], rows[1])
    for k, name in enumerate(("index", "face", "t", "shade")):
        a, b = got[k][sl], ref[k][sl]
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, "%s: %s differs at %d pixels, first %d: %s vs %s" % (
            what, name, bad.size, bad[0], a[bad[0]], b[bad[0]])


@pytest.fixture(scope="module")
def terrain(ctx):
    cfg = W.TERRAIN
    n, w, h = cfg["n"], cfg["width"], cfg["height"]
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, n, cfg["seed"])
    cp = W.terrain_camera()
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, n, cfg["seed"])
    osc = O.brickmap_scene(n, slots, masks, vox)
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading(W.LIGHT), threads=16)
    yield sc, cam, ref
    sc.destroy()


@pytest.mark.parametrize("variant,per_cu", [(8, 20), (8, 16), (8, 0), (9, 16), (5, 16), (5, 24),
                                            (5, 0), (6, 16), (2, 0), (0, 0)])
def test_heightfield_timed_shapes_full_frame(ctx, terrain, variant, per_cu):
    """the shapes bench.py pins for the 1080p frame, counters off, on the
    library stream and on a second stream (two frames in flight)"""
    sc, cam, ref = terrain
    cus = ctx.device_info().compute_units
    tiles = {0: (16, 16), 2: (8, 8), 5: (8, 8), 6: (16, 4), 8: (8, 8), 9: (16, 4)}
    _pin(ctx, sc, cam, L.RAY_SHADOW, variant, per_cu * cus, tiles[variant])
    w, h = cam.width, cam.height
    tg1, tg2 = R.RayTargets(ctx, w, h), R.RayTargets(ctx, w, h)
    _poison(tg1)
    _poison(tg2)
    s2 = C.c_void_p()
    L.check(ctx.lib.meso_stream_create(ctx.h, C.byref(s2)))
    try:
        sh = R.shading(W.LIGHT)
        R.raymarch(ctx, sc, cam, L.RAY_SHADOW, sh, tg1)
        R.raymarch(ctx, sc, cam, L.RAY_SHADOW, sh, tg2, s2.value)
        ctx.sync()
        L.check(ctx.lib.meso_stream_sync(ctx.h, s2))
    finally:
        L.check(ctx.lib.meso_stream_destroy(ctx.h, s2))
    _equal(tg1.results(), ref, "variant %d, %d per CU" % (variant, per_cu))
    _equal(tg2.results(), ref, "variant %d, %d per CU, second stream" % (variant, per_cu))


def _brick_slot_two_level(top, node_slots, bx, by, bz, nt):
    node = top[(bx >> 4) + nt * ((by >> 4) + nt * (bz >> 4))]
    if node == EMPTY:
        return EMPTY
    return node_slots[int(node) * 4096 + (bx & 15) + 16 * (by & 15) + 256 * (bz & 15)]


def _check_brick_sample(kind, gen, n, seed, a, rng, n_alloc=600, n_any=600):
    """bricks of a downloaded scene against the oracle's per-brick generator:
    a sample of the allocated bricks and a sample of all brick cells"""
    nb = n // 8
    if kind == L.SCENE_TWO_LEVEL_U16:
        nt = n // 128
        top = a["top"]
        node_at = np.full(int(top[top != EMPTY].max()) + 1, -1, np.int64)
        cells = np.nonzero(top != EMPTY)[0]
        node_at[top[cells]] = cells
        alloc = np.nonzero(a["slots"] != EMPTY)[0]

        def coords(i):
            node, loc = divmod(int(i), 4096)
            tc = int(node_at[node])
            tx, ty, tz = tc % nt, (tc // nt) % nt, tc // (nt * nt)
            return (16 * tx + loc % 16, 16 * ty + (loc // 16) % 16, 16 * tz + loc // 256)

        def slot_of(bx, by, bz):
            return _brick_slot_two_level(top, a["slots"], bx, by, bz, nt)
    else:
        alloc = np.nonzero(a["slots"] != EMPTY)[0]

        def coords(i):
            i = int(i)
            return i % nb, (i // nb) % nb, i // (nb * nb)

        def slot_of(bx, by, bz):
            return a["slots"][bx + nb * (by + nb * bz)]

    picks = [coords(i) for i in rng.choice(alloc, size=min(n_alloc, alloc.size), replace=False)]
    picks += [tuple(int(v) for v in rng.integers(0, nb, 3)) for _ in range(n_any)]
    solid_seen = 0
    for bx, by, bz in picks:
        s, vox, mask = O.gen_brick(gen, n, seed, bx, by, bz)
        slot = slot_of(bx, by, bz)
        if s == 0:
            assert slot == EMPTY, ("brick should be empty", bx, by, bz)
            continue
        solid_seen += 1
        assert slot != EMPTY, ("brick missing", bx, by, bz)
        assert np.array_equal(a["voxels"][slot], vox), ("voxels", bx, by, bz)
        assert np.array_equal(a["masks"][slot], mask), ("mask", bx, by, bz)
    assert solid_seen >= min(n_alloc, alloc.size)


def test_caves_2048_two_level_bands_bitexact(ctx):
    cfg = W.CAVES
    n, w, h, seed = cfg["n"], cfg["width"], cfg["height"], cfg["seed"]
    print("2048^3 caves: generate", flush=True)
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, seed)
    try:
        info = sc.info()
        assert info.n_bricks > 5_000_000 and info.n_nodes > 0
        cp = W.caves_camera(n, w, h)
        cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
        sh = R.shading(W.LIGHT, 5, cfg["ao_samples"], cfg["ao_radius"])
        flags = L.RAY_SHADOW | L.RAY_AO
        tg = R.RayTargets(ctx, w, h, index_bytes=8)
        _poison(tg)
        # the shape the frame loop times: the autotuner's pick for this key
        t = L.TuneResult()
        t.size = C.sizeof(t)
        print("2048^3 caves: autotune the 4K frame", flush=True)
        L.check(ctx.lib.meso_autotune_raymarch(ctx.h, sc.h, L.byref(cam), flags, L.byref(sh),
                                               L.byref(tg.out), L.byref(t)))
        _poison(tg)
        R.raymarch(ctx, sc, cam, flags, sh, tg)
        ctx.sync()
        got = tg.results()
        print("2048^3 caves: tuned variant %d grid %d; download the volume" % (t.variant, t.grid),
              flush=True)
        a = sc.arrays()
    finally:
        sc.destroy()
    _check_brick_sample(L.SCENE_TWO_LEVEL_U16, O.GEN_CAVES, n, seed, a,
                        np.random.default_rng(2048))
    osc = O.two_level_scene(n, a["top"], a["slots"], a["masks"], a["voxels"])
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
    osh = O.shading(W.LIGHT, 5, cfg["ao_samples"], cfg["ao_radius"])
    big = 0
    for r0, r1 in ((0, 16), (1072, 1096), (2144, 2160)):
        print("2048^3 caves: oracle rows %d-%d" % (r0, r1), flush=True)
        ref = O.raymarch(osc, ocam, flags, osh, row0=r0, row1=r1, threads=16)
        _equal(got, ref, "2048^3 caves rows %d-%d" % (r0, r1), (r0 * w, r1 * w))
        band = got[0][r0 * w:r1 * w]
        big += int((band >= (1 << 31)).sum())
        assert (band >= 0).mean() > 0.5
    assert big > 0, "no hit index at or above 2^31 in the bands"


def test_served_requests_on_256_to_1024_scenes_match_oracle(ctx):
    sh = R.shading(W.LIGHT)
    edges = (256, 520, 1024)
    scenes = [ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, e, 60 + i)
              for i, e in enumerate(edges)]
    try:
        srv = R.Server(ctx, 256 << 20)
        got, cams = [], []
        rng = np.random.default_rng(1024)
        for s, e in zip(scenes, edges):
            v = rng.normal(size=3)
            v[1] = abs(v[1]) * 0.8 + 0.2
            v /= np.linalg.norm(v)
            c = np.array([e / 2, e * 0.15, e / 2])
            eye = c + 1.5 * e * np.sqrt(3.0) * v
            w, hh = (640, 360) if e < 1024 else (1280, 720)
            cams.append((np.float32(eye), np.float32(c), w, hh))
            got.append(srv.submit(s, R.camera(np.float32(eye), np.float32(c), (0, 1, 0), 60.0, w, hh),
                                  L.RAY_SHADOW, sh))
        res = []
        for tk, (_, _, w, hh) in zip(got, cams):
            srv.wait(tk, 60_000_000)
            res.append(srv.take(tk, w, hh))
        srv.stop()
        for i, (s, e) in enumerate(zip(scenes, edges)):
            a = s.arrays()
            _check_brick_sample(L.SCENE_BRICKMAP_U16, O.GEN_TERRAIN_CAVES, e, 60 + i, a,
                                np.random.default_rng(e), 300, 300)
            eye, c, w, hh = cams[i]
            ocam = O.camera(list(eye), list(c), [0.0, 1.0, 0.0], 60.0, w, hh)
            ref = O.raymarch(O.brickmap_scene(e, a["slots"], a["masks"], a["voxels"]), ocam,
                             O.RAY_SHADOW, O.shading(W.LIGHT), threads=16)
            _equal(res[i], ref, "served request on the %d^3 scene" % e)
            assert (res[i][0] >= 0).sum() > 5000  # not a frame of sky
    finally:
        for s in scenes:
            s.destroy()
