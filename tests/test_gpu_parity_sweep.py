"""A seeded sweep of cameras against the oracle, through the C-ABI.

Each scene kind (dense u8, brick maps with an even and an odd edge, two-level)
is rendered from cameras drawn from a fixed seed: eyes anywhere in a box three
times the scene's size (the first two inside the scene), targets inside the
scene, random up vectors and fields of view, and every ray flag combination.
Hit index, face, t and shade must equal the oracle's bit for bit, and the
step counters must be equal too.
"""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu

FLAGS = [0, O.RAY_SHADOW, O.RAY_AO, O.RAY_SHADOW | O.RAY_AO]


def _cameras(n, seed, count):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        # the first two eyes inside the scene volume, the rest anywhere near it
        eye = rng.uniform(0.1 * n, 0.9 * n, 3) if len(out) < 2 else rng.uniform(-n, 2 * n, 3)
        target = rng.uniform(0, n, 3)
        up = rng.normal(size=3)
        fwd = target - eye
        if np.linalg.norm(fwd) < 1.0:
            continue
        c = np.cross(fwd / np.linalg.norm(fwd), up / np.linalg.norm(up))
        if np.linalg.norm(c) < 0.2:
            continue
        out.append(dict(eye=[float(v) for v in eye], target=[float(v) for v in target],
                        up=[float(v) for v in up], vfov=float(rng.uniform(30.0, 100.0)),
                        flags=FLAGS[len(out) % len(FLAGS)]))
    return out


def _scene(ctx, kind):
    if kind == "dense":
        sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 3, 8)
        return sc, O.dense_scene(O.spheres_dense(3, 128, 8)), 128
    if kind == "two":
        n, seed = 256, 5
        sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, seed)
        slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, seed)
        top, slots, masks, vox = O.two_level_from_bricks(n, slots, masks, vox)
        return sc, O.two_level_scene(n, top, slots, masks, vox), n
    gen, n, seed = {"height": (O.GEN_HEIGHTFIELD, 256, 2),
                    "odd": (O.GEN_TERRAIN_CAVES, 200, 8)}[kind]
    sc = ctx.scene_generate(gen, L.SCENE_BRICKMAP_U16, n, seed)
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    return sc, O.brickmap_scene(n, slots, masks, vox), n


@pytest.mark.parametrize("kind,seed", [("dense", 11), ("height", 12), ("odd", 13), ("two", 14)])
def test_random_cameras_match_oracle(ctx, kind, seed):
    sc, osc, n = _scene(ctx, kind)
    w, h = 64, 48
    tg = R.RayTargets(ctx, w, h, counters=True)
    hits = 0
    for i, cp in enumerate(_cameras(n, seed, 8)):
        flags = cp["flags"]
        sh_seed = seed * 100 + i
        cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
        ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h)
        tg.counters.zero()
        R.raymarch(ctx, sc, cam, flags | L.RAY_COUNT, R.shading(W.LIGHT, sh_seed, 4, 12.0), tg)
        ctx.sync()
        idx, face, t, shade = tg.results()
        ridx, rface, rt, rshade, cnt = O.raymarch(osc, ocam, flags,
                                                  O.shading(W.LIGHT, sh_seed, 4, 12.0), threads=8)
        what = "%s camera %d flags %d" % (kind, i, flags)
        assert np.array_equal(idx, ridx), what
        assert np.array_equal(face, rface), what
        assert np.array_equal(t.view(np.uint32), rt.view(np.uint32)), what
        assert np.array_equal(shade.view(np.uint32), rshade.view(np.uint32)), what
        got = tg.counts()
        for k, v in cnt.as_dict().items():
            assert got[k] == v, (what, k)
        hits += int((idx >= 0).sum())
    assert hits > 1000, "the sweep saw too few hits to test anything"
    sc.destroy()


def _axis_cameras(n):
    """Cameras looking straight along each axis, with up along another axis:
    a W x 1 image has primary rays with one component exactly 0, a 1 x H
    image another."""
    c = n * 0.5 + 0.25
    out = []
    for a in range(3):
        for sgn in (1.0, -1.0):
            eye = [c, c, c]
            eye[a] -= sgn * 0.45 * n
            target = [c, c, c]
            up = [0.0, 0.0, 0.0]
            up[(a + 1) % 3] = 1.0
            out.append((eye, target, up))
    return out


@pytest.mark.parametrize("kind", ["dense", "height", "two"])
def test_axis_aligned_rays_match_oracle(ctx, kind):
    """Rays with direction components exactly 0 (primary rays of axis-aligned
    cameras in W x 1 and 1 x H images, shadow rays toward an axis-aligned
    light) against the oracle, with and without the step counters."""
    sc, osc, n = _scene(ctx, kind)
    flags = O.RAY_SHADOW | O.RAY_AO
    zero_primary = 0
    hits = 0
    for i, (eye, target, up) in enumerate(_axis_cameras(n)):
        for w, h in ((64, 1), (1, 48), (24, 18)):
            light = (0.0, 1.0, 0.0) if i % 2 == 0 else (1.0, 0.0, 0.0)
            cam = R.camera(eye, target, up, 60.0, w, h)
            ocam = O.camera(eye, target, up, 60.0, w, h)
            for py in range(h):
                for px in range(w):
                    zero_primary += int((O.primary_ray(ocam, px, py)[1] == 0.0).any())
            ref = O.raymarch(osc, ocam, flags, O.shading(light, 7 + i, 4, 12.0), threads=8)
            for count in (True, False):
                tg = R.RayTargets(ctx, w, h, counters=count)
                if count:
                    tg.counters.zero()
                R.raymarch(ctx, sc, cam, flags | (L.RAY_COUNT if count else 0),
                           R.shading(light, 7 + i, 4, 12.0), tg)
                ctx.sync()
                idx, face, t, shade = tg.results()
                what = "%s camera %d %dx%d count %d" % (kind, i, w, h, count)
                assert np.array_equal(idx, ref[0]), what
                assert np.array_equal(face, ref[1]), what
                assert np.array_equal(t.view(np.uint32), ref[2].view(np.uint32)), what
                assert np.array_equal(shade.view(np.uint32), ref[3].view(np.uint32)), what
                if count:
                    got = tg.counts()
                    for k, v in ref[4].as_dict().items():
                        assert got[k] == v, (what, k)
            hits += int((ref[0] >= 0).sum())
    assert zero_primary > 100, "no primary ray had a zero component"
    assert hits > 200, "too few hits to test anything"
    sc.destroy()
