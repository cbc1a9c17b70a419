"""Pins of the oracle's DDA traversal (CPU only).

Brute force: the FP64 slab intersection of the ray with every solid voxel,
smallest entry t, entry axis by the largest slab entry (ties x, then y, then
z).  The FP32 DDA must pick the same voxel and face on every ray where no
other candidate is within a relative 1e-4 of the first hit (near-ties are
where FP32 and FP64 may legitimately disagree), and its t must be within 1e-5
relative of the FP64 entry-plane t of the voxel it chose, on every hit.
"""
import numpy as np
import pytest

import oracle as O


def brute_first_hit(grid, o, d):
    """grid [z][y][x] bool; returns (t, index, face, second_t) in FP64."""
    n = grid.shape[0]
    zs, ys, xs = np.nonzero(grid)
    if len(xs) == 0:
        return np.inf, -1, 0xFF, np.inf
    lo = np.stack([xs, ys, zs], 1).astype(np.float64)
    hi = lo + 1.0
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (lo - o) / d
        t1 = (hi - o) / d
    tn = np.where(d == 0, np.where((o >= lo) & (o < hi), -np.inf, np.inf), np.minimum(t0, t1))
    tf = np.where(d == 0, np.where((o >= lo) & (o < hi), np.inf, -np.inf), np.maximum(t0, t1))
    tenter = tn.max(1)
    texit = tf.min(1)
    ok = (tenter <= texit) & (texit >= 0)
    if not ok.any():
        return np.inf, -1, 0xFF, np.inf
    te = np.where(ok, np.maximum(tenter, 0.0), np.inf)
    order = np.argsort(te, kind="stable")
    k = order[0]
    second = te[order[1]] if len(order) > 1 else np.inf
    if tenter[k] < 0:
        face = O.FACE_INSIDE
    else:
        axis = int(np.argmax(tn[k]))  # first max: x before y before z
        pos = d[axis] < 0  # travelling -axis enters through the + face
        face = {2: 0, 1: 2, 0: 4}[axis] + (0 if pos else 1)
    idx = int(xs[k] + n * ys[k] + n * n * zs[k])
    return float(te[k]), idx, face, float(second)


def random_grid(rng, n, fill):
    return rng.random((n, n, n)) < fill


def random_rays(rng, n, count, inside_frac=0.2):
    rays = []
    for _ in range(count):
        if rng.random() < inside_frac:
            o = rng.uniform(0, n, 3)
        else:
            v = rng.normal(size=3)
            o = n / 2 + v / np.linalg.norm(v) * n * rng.uniform(0.9, 2.0)
        tgt = rng.uniform(-0.2 * n, 1.2 * n, 3)
        d = tgt - o
        d = d / np.linalg.norm(d)
        if rng.random() < 0.1:  # axis-parallel and axis-plane rays
            d[rng.integers(3)] = 0.0
            if np.linalg.norm(d) == 0:
                d[0] = 1.0
            d = d / np.linalg.norm(d)
        rays.append((o.astype(np.float32), d.astype(np.float32)))
    return rays


def plane_t64(o, d, idx, face, n):
    x, y, z = idx % n, (idx // n) % n, idx // (n * n)
    cell = (x, y, z)
    axis = 2 if face < 2 else (1 if face < 4 else 0)
    plane = cell[axis] + (1 if face % 2 == 0 else 0)
    return (plane - float(o[axis])) / float(d[axis])


def check_scene(scene, grid, rays, n):
    near = mism = 0
    for o, d in rays:
        h, _ = O.trace(scene, o, d)
        t64, idx64, face64, second = brute_first_hit(grid, o, d)
        if idx64 < 0:
            assert h.index < 0 or abs(h.t - t64) < 1e-3, (o, d, h.index)
            if h.index >= 0:
                near += 1
            continue
        if h.index >= 0 and 0 <= h.face < 6:
            # t of the chosen voxel's entry plane, FP64
            assert abs(h.t - plane_t64(o, d, h.index, h.face, n)) <= 1e-5 * max(1.0, abs(h.t)), (o, d)
        tie = (second - t64) <= 1e-4 * max(1.0, t64)
        if tie:
            near += 1
            continue
        if (h.index, h.face) != (idx64, face64):
            # a ray grazing a voxel edge: the entry point lies within 1e-4 of two planes
            p = o.astype(np.float64) + d.astype(np.float64) * t64
            frac = np.abs(p - np.round(p))
            if (frac < 1e-4).sum() >= 2:
                near += 1
                continue
            mism += 1
            raise AssertionError("ray %s %s: dda (%d,%d,%g) brute (%d,%d,%g)" % (
                o, d, h.index, h.face, h.t, idx64, face64, t64))
    return near, mism


@pytest.mark.parametrize("n,fill,seed", [(8, 0.05, 1), (12, 0.02, 2), (16, 0.01, 3), (16, 0.2, 4)])
def test_dense_dda_matches_brute_force(n, fill, seed):
    rng = np.random.default_rng(seed)
    g = random_grid(rng, n, fill)
    sc = O.dense_scene(g.astype(np.uint8) * 3)
    near, _ = check_scene(sc, g, random_rays(rng, n, 600), n)
    assert near < 60


@pytest.mark.parametrize("n,seed", [(16, 5), (32, 6)])
def test_brickmap_dda_matches_brute_force(n, seed):
    rng = np.random.default_rng(seed)
    g = np.zeros((n, n, n), bool)
    # clustered occupancy so that whole bricks are empty
    for _ in range(4):
        c = rng.integers(0, n, 3)
        r = rng.integers(1, 4)
        g[max(0, c[2] - r):c[2] + r, max(0, c[1] - r):c[1] + r, max(0, c[0] - r):c[0] + r] = True
    slots, masks, vox = O.brickmap_from_dense(g.astype(np.uint16) * 7)
    sc = O.brickmap_scene(n, slots, masks, vox)
    assert (slots == O.EMPTY_SLOT).any()
    near, _ = check_scene(sc, g, random_rays(rng, n, 600), n)
    assert near < 60


def test_two_level_dda_matches_brute_force():
    n = 256
    rng = np.random.default_rng(9)
    g = np.zeros((n, n, n), bool)
    for _ in range(6):
        c = rng.integers(0, n, 3)
        r = rng.integers(2, 10)
        g[max(0, c[2] - r):c[2] + r, max(0, c[1] - r):c[1] + r, max(0, c[0] - r):c[0] + r] = True
    slots, masks, vox = O.brickmap_from_dense(g.astype(np.uint16) * 2)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    assert (top == O.EMPTY_SLOT).any() and (top != O.EMPTY_SLOT).any()
    sc = O.two_level_scene(n, top, ns, m2, v2)
    rays = random_rays(rng, n, 250)
    # aim half of the rays at solid voxels so most rays hit
    zs, ys, xs = np.nonzero(g)
    for k in range(250):
        i = rng.integers(len(xs))
        tgt = np.array([xs[i], ys[i], zs[i]]) + 0.5
        o = tgt + rng.normal(size=3) * n
        d = (tgt - o) / np.linalg.norm(tgt - o)
        rays.append((o.astype(np.float32), d.astype(np.float32)))
    near, _ = check_scene(sc, g, rays, n)
    assert near < 50


def test_two_level_equals_brickmap_counts_and_hits():
    n = 256
    slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, 5)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    bm = O.brickmap_scene(n, slots, masks, vox)
    tl = O.two_level_scene(n, top, ns, m2, v2)
    cam = O.camera([-40.0, 150.0, -30.0], [128.0, 128.0, 128.0], [0.0, 1.0, 0.0], 60.0, 96, 64)
    a = O.raymarch(bm, cam, O.RAY_SHADOW, O.shading())
    b = O.raymarch(tl, cam, O.RAY_SHADOW, O.shading())
    # node skipping changes the path through empty space only; where both
    # traversals agree on the voxel they agree on every output bit
    same = a[0] == b[0]
    assert same.mean() > 0.99
    for k in range(1, 4):
        assert np.array_equal(a[k][same].view(np.uint8), b[k][same].view(np.uint8))
    assert b[4].n_top > 0 and b[4].n_node > 0


def test_closed_form_single_voxel():
    n = 16
    g = np.zeros((n, n, n), np.uint8)
    g[5, 6, 7] = 1  # voxel (7,6,5)
    sc = O.dense_scene(g)
    rng = np.random.default_rng(1)
    for _ in range(200):
        tgt = np.array([7, 6, 5]) + rng.uniform(0.05, 0.95, 3)
        o = tgt + rng.normal(size=3) * 20
        d = (tgt - o) / np.linalg.norm(tgt - o)
        o, d = o.astype(np.float32), d.astype(np.float32)
        h, _ = O.trace(sc, o, d)
        lo = np.array([7, 6, 5], np.float64)
        od, dd = o.astype(np.float64), d.astype(np.float64)
        t0, t1 = (lo - od) / dd, (lo + 1 - od) / dd
        tn = np.minimum(t0, t1)
        assert h.index == 7 + 16 * 6 + 256 * 5
        assert abs(h.t - tn.max()) <= 1e-5 * tn.max()


def test_closed_form_axis_aligned_box_face():
    n = 32
    g = np.zeros((n, n, n), np.uint8)
    g[10:20, 8:14, 4:30] = 1  # z 10..19, y 8..13, x 4..29
    sc = O.dense_scene(g)
    lo = np.array([4, 8, 10], np.float64)
    hi = np.array([30, 14, 20], np.float64)
    rng = np.random.default_rng(2)
    for _ in range(300):
        tgt = lo + rng.uniform(0.01, 0.99, 3) * (hi - lo)
        o = tgt + rng.normal(size=3) * 40
        d = (tgt - o) / np.linalg.norm(tgt - o)
        o, d = o.astype(np.float32), d.astype(np.float32)
        od, dd = o.astype(np.float64), d.astype(np.float64)
        tn = np.minimum((lo - od) / dd, (hi - od) / dd)
        if np.sort(tn)[-1] - np.sort(tn)[-2] < 1e-3:
            continue  # edge hit
        if (od >= lo).all() and (od < hi).all():
            continue
        axis = int(np.argmax(tn))
        face = {2: 0, 1: 2, 0: 4}[axis] + (0 if dd[axis] < 0 else 1)
        h, _ = O.trace(sc, o, d)
        assert h.face == face
        assert abs(h.t - tn.max()) <= 1e-5 * tn.max()


def test_axis_parallel_rays_distance():
    n = 16
    g = np.zeros((n, n, n), np.uint8)
    g[3, 4, 9] = 1  # x=9, y=4, z=3
    sc = O.dense_scene(g)
    h, cnt = O.trace(sc, np.array([0.5, 4.5, 3.5], np.float32), np.array([1, 0, 0], np.float32))
    assert h.index == 9 + 16 * 4 + 256 * 3 and h.face == 5 and h.t == 8.5
    assert cnt.n_vox == 10  # cells x = 0..9
    h, cnt = O.trace(sc, np.array([9.5, 4.5, -10.0], np.float32), np.array([0, 0, 1], np.float32))
    assert h.face == 1 and h.t == 13.0
    h, cnt = O.trace(sc, np.array([2.5, 2.5, -10.0], np.float32), np.array([0, 0, 1], np.float32))
    assert h.index == -1 and h.face == 0xFF and np.isinf(h.t) and cnt.n_vox == n


def test_brick_walk_counts_by_hand():
    """One brick holding one solid voxel (3,3,3); a ray along +x at y = z = 3.5.
    The walk enters the brick (one slot read, one block word), steps over
    the empty 2^3 block of x 0..1 at once, then tests voxels x = 2 and x = 3
    at the voxel level (two mask-word reads) and hits x = 3."""
    n = 16
    slots = np.full(8, 0xFFFFFFFF, np.uint32)
    slots[0] = 0
    masks = np.zeros((1, 8), np.uint64)
    masks[0, 3] = np.uint64(1) << np.uint64(3 + 8 * 3)
    vox = np.zeros((1, 512), np.uint16)
    vox[0, 3 + 8 * 3 + 64 * 3] = 5
    sc = O.brickmap_scene(n, slots, masks, vox)
    h, cnt = O.trace(sc, np.array([-4.0, 3.5, 3.5], np.float32), np.array([1, 0, 0], np.float32))
    assert h.index == 3 + 16 * 3 + 256 * 3 and h.face == 5 and h.t == 7.0 and h.material == 5
    assert (cnt.n_brickslot, cnt.n_brickmask, cnt.n_vox, cnt.n_iter) == (1, 1, 2, 3)
    # the same ray one row higher (y = 4.5) stays in brick 0's 4^3 block of
    # y 4..7, which is empty: one step across the brick, no voxel test
    h, cnt = O.trace(sc, np.array([-4.0, 4.5, 3.5], np.float32), np.array([1, 0, 0], np.float32))
    assert h.index == -1 and (cnt.n_brickmask, cnt.n_vox) == (1, 0)


def test_raymarch_invariants_sphere_grid():
    g = O.spheres_dense(1, 128, 8)
    sc = O.dense_scene(g)
    import math
    c = np.array([64, 64, 64.0])
    dv = np.array([1, 0.8, 0.6]) / np.linalg.norm([1, 0.8, 0.6])
    eye = (c + 1.5 * 128 * math.sqrt(3) * dv).astype(np.float32)
    cam = O.camera(eye, c.astype(np.float32), (0, 0, 1), 60.0, 128, 128)
    idx, face, t, shade, cnt = O.raymarch(sc, cam, O.RAY_SHADOW, O.shading())
    flat = g.ravel()
    hit = idx >= 0
    assert hit.any()
    assert (flat[idx[hit]] > 0).all()  # the hit voxel is solid
    # the neighbour across the hit face (where the ray came from) is empty
    n = 128
    x, y, z = idx[hit] % n, (idx[hit] // n) % n, idx[hit] // (n * n)
    f = face[hit].astype(np.int64)
    nb = np.stack([x, y, z], 1)
    axis = np.where(f < 2, 2, np.where(f < 4, 1, 0))
    sgn = np.where(f % 2 == 0, 1, -1)
    nb[np.arange(len(f)), axis] += sgn
    inside = ((nb >= 0) & (nb < n)).all(1)
    nbi = nb[inside]
    assert (flat[nbi[:, 0] + n * nbi[:, 1] + n * n * nbi[:, 2]] == 0).all()
    assert np.isinf(t[~hit]).all() and (face[~hit] == 0xFF).all() and (shade[~hit] == 0).all()
    # shade = (0.25 + 0.75 * lambert * visible) on lit faces, 0.25 in shadow
    L = np.array([0.4, 1.0, 0.3]) / np.linalg.norm([0.4, 1.0, 0.3])
    lam = np.where(axis == 2, L[2], np.where(axis == 1, L[1], L[0])) * sgn
    lam = np.maximum(lam, 0)
    s = shade[hit].astype(np.float64)
    lit = s > 0.2500001
    assert np.allclose(s[lit], 0.25 + 0.75 * lam[lit], rtol=1e-5)
    assert np.allclose(s[~lit], 0.25)


def test_ao_directions_cosine_hemisphere():
    # Malley's method: z = sqrt(1 - r^2) with (a,b) uniform in the disk gives
    # E[cos] = 2/3 and unit length; the normal component is always positive
    zs = []
    for p in range(4000):
        d = O.ao_direction(7, p, p % 4, 2)  # +Y face
        assert abs(np.linalg.norm(d.astype(np.float64)) - 1) < 1e-6
        assert d[1] > 0
        zs.append(d[1])
    assert abs(np.mean(zs) - 2 / 3) < 0.02
    d = O.ao_direction(7, 1, 0, 5)  # -X face
    assert d[0] < 0
