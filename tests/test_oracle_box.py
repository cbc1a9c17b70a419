"""Pins of the oracle's solid box (CPU only): rays are clipped to the box of
the allocated bricks and walks end when they step out of it.

- The box equals a numpy scan of the brick slots (brick map and two-level),
  and is empty for a scene without bricks.
- Clipping changes no output: on whole frames with shadow and AO rays, the
  index, face, t and shade are bit-identical to the same walk over the whole
  grid (box widened to [0, n)), with fewer traversal steps.
"""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import workloads as W


def _numpy_box(slots, n):
    nb = n // 8
    s = slots.reshape(nb, nb, nb)  # (z, y, x)
    z, y, x = np.nonzero(s != O.EMPTY_SLOT)
    if len(x) == 0:
        return (0, 0, 0), (0, 0, 0)
    return ((8 * x.min(), 8 * y.min(), 8 * z.min()),
            (8 * (x.max() + 1), 8 * (y.max() + 1), 8 * (z.max() + 1)))


@pytest.mark.parametrize("gen,n,seed", [(O.GEN_HEIGHTFIELD, 256, 2), (O.GEN_CAVES, 256, 5),
                                        (O.GEN_TERRAIN_CAVES, 200, 8)])
def test_box_is_the_allocated_bricks(gen, n, seed):
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    want = _numpy_box(slots, n)
    assert O.brickmap_scene(n, slots, masks, vox).box == want
    if n % 128 == 0:
        top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
        assert O.two_level_scene(n, top, ns, m2, v2).box == want


def test_empty_scene_box_misses_everything():
    g = np.zeros((64, 64, 64), np.uint16)
    slots, masks, vox = O.brickmap_from_dense(g)
    sc = O.brickmap_scene(64, slots, masks, vox)
    assert sc.box == ((0, 0, 0), (0, 0, 0))
    cam = O.camera([100.0, 80.0, 90.0], [32.0, 32.0, 32.0], [0.0, 1.0, 0.0], 60.0, 32, 24)
    idx, face, t, shade, _ = O.raymarch(sc, cam, O.RAY_SHADOW, O.shading(W.LIGHT))
    assert (idx == -1).all() and (face == O.FACE_MISS).all()


def _widen(sc):
    for a in range(3):
        sc.s.box_lo[a] = 0
        sc.s.box_hi[a] = sc.n


@pytest.mark.parametrize("gen,n,seed,eye,target,flags", [
    (O.GEN_HEIGHTFIELD, 256, 2, (-30.0, 130.0, 128.0), (150.0, 15.0, 128.0), O.RAY_SHADOW),
    (O.GEN_HEIGHTFIELD, 256, 2, (128.0, 60.0, 128.0), (200.0, 20.0, 40.0),
     O.RAY_SHADOW | O.RAY_AO),
    (O.GEN_TERRAIN_CAVES, 200, 8, (100.0, 190.0, -40.0), (100.0, 40.0, 100.0),
     O.RAY_SHADOW | O.RAY_AO)])
def test_clipping_changes_no_output(gen, n, seed, eye, target, flags):
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    sc = O.brickmap_scene(n, slots, masks, vox)
    lo, hi = sc.box
    assert lo != (0, 0, 0) or hi != (n, n, n), "the box should be smaller than the grid"
    cam = O.camera(list(eye), list(target), [0.0, 1.0, 0.0], 70.0, 160, 120)
    sh = O.shading(W.LIGHT, 3, 4, 12.0)
    a = O.raymarch(sc, cam, flags, sh, threads=8)
    _widen(sc)
    b = O.raymarch(sc, cam, flags, sh, threads=8)
    for x, y in zip(a[:4], b[:4]):
        assert x.tobytes() == y.tobytes()
    ca, cb = a[4].as_dict(), b[4].as_dict()
    assert (a[0] >= 0).mean() > 0.2
    assert ca["n_iter"] < cb["n_iter"] and ca["hits"] == cb["hits"]
