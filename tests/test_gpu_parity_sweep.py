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
