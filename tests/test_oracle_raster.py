"""Pins of the oracle's quad visibility (CPU only).

- A float64 ray-quad brute force over every pixel of hand-placed quads
  (stacked, back-facing, grazing, partly behind the eye): the nearest
  front-facing quad per pixel, t within 1e-5 relative; pixels whose float64
  hit lies within 1e-4 of a quad edge, or whose two nearest quads are within
  1e-4 in t, are counted and bounded, not compared.
- Whole frames of fully resident scenes against the oracle's DDA (primary
  rays, flags 0): the voxel, face, t (bit for bit) and shade agree on at
  least 99.5% of the pixels, and every pixel where they differ has its
  float64 DDA hit within 1e-3 of a voxel edge.
"""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import workloads as W

LIGHT = O.shading(W.LIGHT).light


def _pack(x, y, z, w, h, f):
    return O.lib().mo_quad_pack(x, y, z, w, h, f)


def _axes(f):
    return (2, 0, 1) if f < 2 else ((1, 0, 2) if f < 4 else (0, 1, 2))


def _brute(cam, quads):
    """float64 nearest front-facing quad per pixel: (quad k or -1, t, near-tie flag)"""
    c = cam
    w, h = c.width, c.height
    px, py = np.meshgrid(np.arange(w), np.arange(h))
    a = ((px + 0.5) * (2.0 / w) - 1.0) * c.sx
    b = (1.0 - (py + 0.5) * (2.0 / h)) * c.sy
    R, U, F = (np.array(list(v), np.float64) for v in (c.right, c.up, c.fwd))
    d = F[None, None, :] + a[..., None] * R + b[..., None] * U
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    o = np.array(list(c.eye), np.float64)
    best = np.full((h, w), np.inf)
    second = np.full((h, w), np.inf)
    which = np.full((h, w), -1)
    edge = np.zeros((h, w), bool)
    for k, (x, y, z, qw, qh, f) in enumerate(quads):
        A, Uax, Vax = _axes(f)
        pos = f % 2 == 0
        cell = (x, y, z)
        P = cell[A] + (1 if pos else 0)
        u0, v0 = cell[Uax], cell[Vax]
        da = d[..., A]
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (P - o[A]) / da
        hu = o[Uax] + d[..., Uax] * t
        hv = o[Vax] + d[..., Vax] * t
        front = (da < 0) if pos else (da > 0)
        inside = (hu >= u0) & (hu < u0 + qw) & (hv >= v0) & (hv < v0 + qh)
        hit = front & (t > 0) & inside
        near_edge = front & (t > 0) & (np.minimum.reduce([np.abs(hu - u0), np.abs(hu - u0 - qw),
                                                          np.abs(hv - v0), np.abs(hv - v0 - qh)])
                                       < 1e-4) & (hu > u0 - 1e-4) & (hu < u0 + qw + 1e-4) \
            & (hv > v0 - 1e-4) & (hv < v0 + qh + 1e-4)
        edge |= near_edge
        tt = np.where(hit, t, np.inf)
        closer = tt < best
        second = np.where(closer, best, np.minimum(second, tt))
        which = np.where(closer, k, which)
        best = np.where(closer, tt, best)
    with np.errstate(invalid="ignore"):
        tie = np.isfinite(second) & (second - best < 1e-4 * np.maximum(best, 1.0))
    return which.ravel(), best.ravel(), (edge | tie).ravel()


def test_hand_placed_quads_against_float64():
    quads = [(4, 6, 10, 20, 12, 0),   # +Z at z = 11
             (10, 10, 20, 5, 5, 0),   # +Z at z = 21, in front of the first
             (0, 0, 25, 32, 32, 1),   # -Z at z = 25: back-facing from above
             (2, 3, 1, 8, 20, 4),     # +X at x = 3, seen at a grazing angle
             (0, 28, 0, 32, 4, 2),    # +Y at y = 29, partly behind the eye plane
             (20, 0, 0, 12, 30, 3)]   # -Y at y = 0: back-facing for every eye above it
    q = np.array([_pack(*v) for v in quads], np.uint32)
    for eye, target in [((16.0, 16.0, 60.0), (16.0, 16.0, 0.0)),
                        ((30.0, 40.0, 40.0), (12.0, 12.0, 10.0)),
                        ((16.0, 29.5, 14.0), (16.0, 5.0, 12.0))]:
        cam = O.camera(list(eye), list(target), [0.0, 0.0, 1.0] if eye[2] < 20 else [0.0, 1.0, 0.0],
                       60.0, 96, 80)
        idx, face, t, shade, keys = O.raster_pool(cam, 32, 1, np.zeros(1, np.uint32),
                                                  np.array([len(q)], np.uint32), q, len(q), LIGHT)
        which, tb, near = _brute(cam, quads)
        got = np.where(keys == np.uint64(0xFFFFFFFFFFFFFFFF), -1,
                       (keys & np.uint64(0xFFFFFFFF)).astype(np.int64))
        ok = ~near
        assert ok.mean() > 0.95
        assert np.array_equal(got[ok], which[ok]), (eye, np.nonzero(got[ok] != which[ok])[0][:5])
        hit = ok & (which >= 0)
        assert hit.sum() > 500
        assert np.all(np.abs(t[hit] - tb[hit]) <= 1e-5 * tb[hit])
        assert np.all(face[hit] == np.array([quads[k][5] for k in which[hit]]))
        miss = ok & (which < 0)
        assert np.all(idx[miss] == -1) and np.all(np.isinf(t[miss])) and np.all(shade[miss] == 0)


def _resident_pool(sc, n):
    nc = -(-n // 32)
    R = int(np.ceil(np.sqrt(3) * nc)) + 1
    p = O.pool_params(forward_radius=R, backward_radius=R, view_angle=120.0, capacity=nc ** 3,
                      quads_per_slot=98304, max_loads=nc ** 3, greedy=True)
    pool = O.OraclePool(sc, p)
    pool.update((nc // 2, nc // 2, nc // 2), (1.0, 0.0, 0.0))
    res = pool.slot_chunk != O.NO_SLOT
    want = sum(O.chunk_nonempty(sc, x, y, z) for z in range(nc) for y in range(nc)
               for x in range(nc))
    assert res.sum() == want
    return pool, nc


def _near_edge(cam, t, face, pix):
    """the float64 hit point of each pixel (at t on a plane of the face's
    axis) lies within 1e-3 of a voxel edge"""
    out = []
    for i in pix:
        o, d = O.primary_ray(cam, int(i % cam.width), int(i // cam.width))
        p = o.astype(np.float64) + d.astype(np.float64) * float(t[i])
        a = 2 if face[i] < 2 else (1 if face[i] < 4 else 0)
        others = [k for k in range(3) if k != a]
        frac = np.abs(p[others] - np.round(p[others]))
        out.append(bool((frac < 1e-3).any()))
    return np.array(out, bool)


@pytest.mark.parametrize("kind,gen,n,seed,eye,target", [
    ("brick", O.GEN_HEIGHTFIELD, 128, 2, (-30.0, 90.0, -20.0), (64.0, 30.0, 64.0)),
    ("two", O.GEN_CAVES, 256, 5, (128.0, 140.0, 120.0), (200.0, 100.0, 180.0)),
    ("two", O.GEN_CAVES, 256, 5, (-40.0, 300.0, -40.0), (128.0, 100.0, 128.0))])
def test_whole_frames_against_the_dda(kind, gen, n, seed, eye, target):
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    if kind == "two":
        top, slots, masks, vox = O.two_level_from_bricks(n, slots, masks, vox)
        sc = O.two_level_scene(n, top, slots, masks, vox)
    else:
        sc = O.brickmap_scene(n, slots, masks, vox)
    pool, nc = _resident_pool(sc, n)
    cam = O.camera(list(eye), list(target), [0.0, 1.0, 0.0], 70.0, 160, 120)
    ri, rf, rt, rs, _ = O.raster_pool(cam, n, nc, pool.slot_chunk, pool.slot_count, pool.quads,
                                      pool.p.quads_per_slot, LIGHT)
    di, df, dt, ds, _ = O.raymarch(sc, cam, 0, O.shading(W.LIGHT), threads=8)
    assert (di >= 0).mean() > 0.2
    assert not (df == 6).any(), "the camera starts inside a voxel"
    same = (ri == di) & (rf == df)
    assert same.mean() >= 0.995, same.mean()
    assert np.array_equal(rt.view(np.uint32)[same], dt.view(np.uint32)[same])
    assert np.array_equal(rs.view(np.uint32)[same], ds.view(np.uint32)[same])
    diff = np.nonzero(~same)[0]
    if len(diff):
        t_any = np.where(di >= 0, dt, rt)
        f_any = np.where(di >= 0, df, rf)
        near = _near_edge(cam, t_any, f_any, diff)
        assert near.all(), diff[~near][:5]
