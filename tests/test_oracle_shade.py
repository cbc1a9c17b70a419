"""Pins of the oracle's shading (CPU only): the shade of every hit pixel,
with shadow rays and with shadow + AO rays, recomputed in FP64 from a brute
force occlusion test against every solid voxel.

shade = (0.25 + 0.75 * lambert * visible) * (0.5 + 0.5 * unoccluded / samples)
with lambert = max(0, n . l) for the hit face's normal n and the unit light
l; visible = 0 when the shadow ray toward l from the hit point meets a solid
voxel; unoccluded = the AO rays (directions from the oracle's Malley sampler,
pinned by test_ao_directions_cosine_hemisphere) that meet no solid voxel
within the AO radius.  Rays that graze a voxel edge or end within 1e-4 of a
voxel boundary are near-ties where FP32 and FP64 may disagree; they are
counted and bounded, not compared.
"""
import numpy as np
import pytest

import oracle as O

LIGHT = np.array([0.4, 1.0, 0.3])


def _boxes(grid):
    zs, ys, xs = np.nonzero(grid)
    lo = np.stack([xs, ys, zs], 1).astype(np.float64)
    return lo, lo + 1.0


def brute_occluded(lo, hi, p, d, tmax, own):
    """(occluded, near_tie) of the segment p + t d, 0 < t < tmax, FP64, over
    the solid voxels other than `own` (the hit voxel the ray leaves: in FP32
    the hit point may sit a rounding step inside it).  A box counts when the
    ray passes through its interior inside the segment.
    Near-ties: a box only grazed (entry and exit within 1e-4), a box entered
    within 1e-4 of tmax, or a box touched at the start point (entered within
    1e-4 of t = 0) when no box is hit clearly further on."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - p) / d
        t1 = (hi - p) / d
    inside = (p >= lo) & (p <= hi)
    tn = np.where(d == 0, np.where(inside, -np.inf, np.inf), np.minimum(t0, t1)).max(1)
    tf = np.where(d == 0, np.where(inside, np.inf, -np.inf), np.maximum(t0, t1)).min(1)
    eps = 1e-4
    mine = (lo == np.asarray(own, np.float64)).all(1)
    tn = np.where(mine, np.inf, tn)
    tf = np.where(mine, -np.inf, tf)
    hit = (tn < tf) & (tf > 1e-7) & (tn < tmax)
    clear = hit & (tn > eps) & (tn < tmax - eps * max(1.0, tmax if np.isfinite(tmax) else 1.0)) \
        & (tf - tn > eps)
    graze = (tf - tn > -eps) & (tf - tn <= eps) & (tf > -eps) & (tn < tmax + eps)
    at_end = (np.abs(tn - tmax) <= eps * max(1.0, tmax)) if np.isfinite(tmax) else \
        np.zeros_like(tn, bool)
    at_start = (np.abs(tn) <= eps) & (tf > eps)
    tie = bool(graze.any() or at_end.any() or (at_start.any() and not clear.any()))
    return bool(hit.any()), tie


def _scene(seed, n=16):
    rng = np.random.default_rng(seed)
    g = np.zeros((n, n, n), bool)
    g[0:2, :, :] = True  # a floor (z = 0, 1) that the boxes cast shadows on
    for _ in range(6):
        c = rng.integers(2, n - 2, 3)
        r = rng.integers(1, 3, 3)
        g[c[2] - r[2]:c[2] + r[2], c[1] - r[1]:c[1] + r[1], c[0] - r[0]:c[0] + r[0]] = True
    return g


@pytest.mark.parametrize("ao", [False, True])
def test_shade_matches_fp64_brute_force(ao):
    n = 16
    grid = _scene(11, n)
    sc = O.dense_scene(grid.astype(np.uint8) * 3)
    lo, hi = _boxes(grid)
    light = LIGHT / np.linalg.norm(LIGHT)
    w, h = 96, 80
    cam = O.camera([26.0, -14.0, 30.0], [8.0, 8.0, 2.0], [0.0, 0.0, 1.0], 60.0, w, h)
    flags = O.RAY_SHADOW | (O.RAY_AO if ao else 0)
    seed, samples, radius = 9, 4, 6.0
    idx, face, t, shade, cnt = O.raymarch(sc, cam, flags, O.shading(tuple(LIGHT), seed, samples,
                                                                     radius))
    checked = ties = shadowed = 0
    for pix in np.nonzero((face < 6) & (idx >= 0))[0]:
        px, py = int(pix % w), int(pix // w)
        o, d = O.primary_ray(cam, px, py)
        f = int(face[pix])
        p = o.astype(np.float64) + d.astype(np.float64) * float(t[pix])
        own = (int(idx[pix]) % n, (int(idx[pix]) // n) % n, int(idx[pix]) // (n * n))
        axis = 2 if f < 2 else (1 if f < 4 else 0)
        sign = 1.0 if f % 2 == 0 else -1.0
        lam = max(0.0, sign * light[axis])
        tie = False
        vis = 1.0
        if lam > 0:
            occ, tv = brute_occluded(lo, hi, p, light, np.inf, own)
            tie |= tv
            vis = 0.0 if occ else 1.0
            shadowed += occ
        aof = 1.0
        if ao:
            unocc = 0
            for k in range(samples):
                dk = O.ao_direction(seed, int(pix), k, f).astype(np.float64)
                occ, tv = brute_occluded(lo, hi, p, dk, radius, own)
                tie |= tv
                unocc += 0 if occ else 1
            aof = 0.5 + 0.5 * unocc / samples
        want = (0.25 + 0.75 * lam * vis) * aof
        if tie:
            ties += 1
            continue
        checked += 1
        assert abs(float(shade[pix]) - want) <= 1e-5 * want, (pix, f, float(shade[pix]), want)
    assert checked > 700 and shadowed > 80, (checked, shadowed)
    assert ties <= 0.05 * (checked + ties), (ties, checked)
