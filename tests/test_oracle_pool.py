"""Pins of the oracle's chunk pool (CPU only).

- Importance (CalculateChunkImportance, Runtimes/Voxel/Chunk/ChunkManagerHelper.h:21-43):
  closed-form values on and off the view axis, inside the 5^3 core and past
  the far distance.
- Desired set (ChunkManagerHelper.h:91-148): with equal forward and backward
  radii it is the integer lattice ball |o|^2 < R^2 plus the 3^3 core for any
  view direction; with a zero backward radius nothing behind the camera but
  the core; against a float64 evaluation of the rule for every offset not
  within 1e-4 of a boundary.
- Updates (FindMin replacement, ChunkPool.h:446-560; load limit and re-queue,
  ChunkManager.h:211-330): over a camera path on a small scene, every update
  is checked against an independent Python ranking of the pre-update state:
  loads = the top candidates in order, evictions = the least important
  resident chunks, the stop rule, the load limit, lookup and slots as inverse
  maps, stamps, and each slot's quads against the chunk mesher.
"""
import math

import numpy as np
import pytest

import oracle as O


def test_importance_closed_forms():
    f = O.pool_forward([0.0, 0.0, 3.0])  # +z, any length
    assert list(f) == [0.0, 0.0, 1.0]
    for o in [(0, 0, 0), (2, -2, 1), (-2, 2, -2), (0, 0, 2)]:
        assert O.chunk_importance(*o, f) == 1.0e6
    for d in (3, 10, 40, 63):
        assert O.chunk_importance(0, 0, d, f) == 64.0 - d          # ahead: angle term 1
        assert O.chunk_importance(0, 0, -d, f) == 0.75 * (64.0 - d)  # behind: floor 0.75
        assert O.chunk_importance(d, 0, 0, f) == 0.75 * (64.0 - d)   # beside: floor 0.75
    assert O.chunk_importance(0, 0, 64, f) == 0.25   # distance floor 0.25
    assert O.chunk_importance(0, 0, 100, f) == 0.25
    assert O.chunk_importance(0, 0, -100, f) == 0.75 * 0.25
    # 3-4-5 offset: dot = 0.8 -> angle term (0.8 - 0.5) * 2 = 0.6 -> floor 0.75;
    # length 5 (exact)
    assert O.chunk_importance(0, 3, 4, f) == np.float32(0.75) * np.float32(59.0)
    # dot = 0.96 (7-24-25): angle term (0.96 - 0.5) * 2 = 0.92
    got = O.chunk_importance(0, 7, 24, f)
    assert abs(got - 0.92 * 39.0) < 1e-4, got


@pytest.mark.parametrize("fwd", [(1, 0, 0), (0, -1, 0), (0.3, 0.5, -0.8), (-1, -1, 1)])
@pytest.mark.parametrize("R", [3, 6, 9])
def test_desired_equal_radii_is_the_lattice_ball(fwd, R):
    f = O.pool_forward(fwd)
    p = O.pool_params(R, R, 90.0, 10, 10, 1)
    got = {o for o in _cube(R) if O.chunk_desired(*o, f, p)}
    want = {o for o in _cube(R) if o[0] ** 2 + o[1] ** 2 + o[2] ** 2 < R * R
            or max(abs(v) for v in o) <= 1}
    assert got == want


def test_desired_zero_backward_radius_keeps_only_the_core_behind():
    f = O.pool_forward([1, 0, 0])
    p = O.pool_params(8, 0, 60.0, 10, 10, 1)
    for d in range(2, 9):
        assert O.chunk_desired(-d, 0, 0, f, p) is False
        assert O.chunk_desired(0, d, 0, f, p) is False   # perpendicular: alpha 0
        assert O.chunk_desired(d, 0, 0, f, p) is (d < 8)  # on the axis: forward radius
    for o in _cube(1):
        assert O.chunk_desired(*o, f, p)


def _desired64(o, f, R, RB, angle):
    """the reference's acceptance rule in float64"""
    ln = math.sqrt(o[0] ** 2 + o[1] ** 2 + o[2] ** 2)
    if ln > R + 1e-6:
        return False, False
    if max(abs(v) for v in o) <= 1:
        return True, False
    thr = max(math.cos(math.radians(angle) * 0.5), 0.01)
    a = max(sum(f[i] * o[i] / ln for i in range(3)), 0.0)
    near = abs(a - thr) < 1e-4
    a = 1.0 if a > thr else a / thr
    a = min(max(a, 0.0), 1.0)
    lim = a * R + (1 - a) * RB
    return ln < lim, near or abs(ln - lim) < 1e-4


@pytest.mark.parametrize("fwd,R,RB,angle", [((1, 0.2, -0.1), 12, 4, 120.0),
                                              ((0, 0, 1), 10, 2, 60.0),
                                              ((-0.4, 0.9, 0.3), 16, 6, 90.0)])
def test_desired_against_float64_rule(fwd, R, RB, angle):
    f = O.pool_forward(fwd)
    f64 = np.asarray(fwd, np.float64) / np.linalg.norm(fwd)
    p = O.pool_params(R, RB, angle, 10, 10, 1)
    checked = 0
    for o in _cube(R):
        want, tie = _desired64(o, f64, R, RB, angle)
        if tie:
            continue
        checked += 1
        assert O.chunk_desired(*o, f, p) == want, o
    assert checked > 0.95 * (2 * R + 1) ** 3


@pytest.mark.parametrize("n", [2, 7, 256])
def test_fibonacci_directions(n):
    """GetFibonacciSphere (VoxelMathHelper.h:48-71): unit vectors, heights
    stepping evenly from +1 to -1, and successive azimuths one golden angle
    pi * (3 - sqrt(5)) apart (mod 2 pi), i.e. pi * (sqrt(5) - 1) forward"""
    d = O.pool_bake_directions(n).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    assert np.allclose(d[:, 1], 1.0 - 2.0 * np.arange(n) / (n - 1), atol=1e-6)
    inner = np.abs(d[:, 1]) < 0.999
    az = np.arctan2(d[:, 2], d[:, 0])
    step = np.mod(np.diff(az), 2 * np.pi)
    both = inner[1:] & inner[:-1]
    assert np.allclose(step[both], np.mod(np.pi * (np.sqrt(5) - 1), 2 * np.pi), atol=1e-4)


def test_snap_is_the_nearest_baked_direction():
    d = O.pool_bake_directions(256)
    rng = np.random.default_rng(4)
    for _ in range(200):
        v = rng.normal(size=3)
        f = O.pool_forward(v)
        got = O.pool_snap(d, f)
        dist = np.linalg.norm(d.astype(np.float64) - f.astype(np.float64), axis=1)
        order = np.argsort(dist)
        if dist[order[1]] - dist[order[0]] < 1e-6:
            continue  # a near-tie between two directions
        assert np.array_equal(got, d[order[0]])


def test_baked_pool_equals_exact_pool_on_the_snapped_view():
    sc, _ = _scene()
    p = O.pool_params(forward_radius=4, backward_radius=2, view_angle=100.0, capacity=48,
                      quads_per_slot=3000, max_loads=12, greedy=True)
    baked = O.OraclePool(sc, p, bake_directions=256)
    exact = O.OraclePool(sc, p)
    dirs = O.pool_bake_directions(256)
    for cam, fwd in [((1, 1, 1), (1, 0.3, 0)), ((3, 2, 2), (0.2, 0.1, 1)),
                     ((5, 2, 5), (-1, 0.4, -0.5))]:
        snapped = O.pool_snap(dirs, O.pool_forward(fwd))
        assert baked.update(cam, fwd)[:3] == exact.update(cam, snapped)[:3]
        assert np.array_equal(baked.lookup, exact.lookup)
        assert np.array_equal(baked.slot_chunk, exact.slot_chunk)


def _cube(R):
    r = range(-R, R + 1)
    return [(x, y, z) for z in r for y in r for x in r]


def _scene():
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, 256, 2)
    return O.brickmap_scene(256, slots, masks, vox), slots


def test_nonempty_matches_brick_slots():
    sc, slots = _scene()
    s = slots.reshape(32, 32, 32)  # (z, y, x) bricks
    for cz in range(8):
        for cy in range(8):
            for cx in range(8):
                want = bool((s[4 * cz:4 * cz + 4, 4 * cy:4 * cy + 4, 4 * cx:4 * cx + 4]
                             != O.EMPTY_SLOT).any())
                assert O.chunk_nonempty(sc, cx, cy, cz) == want
    assert not O.chunk_nonempty(sc, 8, 0, 0) and not O.chunk_nonempty(sc, -1, 0, 0)


def _ranked_candidates(pool, cam, f):
    nc, p = pool.nc, pool.p
    out = []
    for o in _cube(p.forward_radius):
        c = (cam[0] + o[0], cam[1] + o[1], cam[2] + o[2])
        if min(c) < 0 or max(c) >= nc:
            continue
        idx = c[0] + nc * (c[1] + nc * c[2])
        if pool.lookup[idx] != O.NO_SLOT or not O.chunk_desired(*o, f, p) \
                or not O.chunk_nonempty(pool.scene, *c):
            continue
        out.append((-O.chunk_importance(*o, f), idx))
    return sorted(out)


def _ranked_victims(pool, cam, f):
    nc = pool.nc
    out = []
    for sl, c in enumerate(pool.slot_chunk):
        if c == O.NO_SLOT:
            out.append((0, -math.inf, sl))
            continue
        x, y, z = int(c) % nc, (int(c) // nc) % nc, int(c) // (nc * nc)
        out.append((1, O.chunk_importance(x - cam[0], y - cam[1], z - cam[2], f), sl))
    return sorted(out)


@pytest.mark.parametrize("greedy", [True, False])
def test_updates_against_an_independent_ranking(greedy):
    sc, _ = _scene()
    p = O.pool_params(forward_radius=4, backward_radius=2, view_angle=100.0, capacity=48,
                      quads_per_slot=3000, max_loads=12, greedy=greedy)
    pool = O.OraclePool(sc, p)
    path = [((1, 1, 1), (1, 0, 0)), ((1, 1, 1), (1, 0, 0)), ((2, 1, 1), (1, 0.2, 0)),
            ((3, 2, 2), (0, 0, 1)), ((5, 2, 5), (-1, 0, -0.5)), ((6, 1, 6), (0, 1, 0)),
            ((6, 1, 6), (0, 1, 0)), ((2, 2, 2), (0.5, -0.5, 0.7)), ((2, 2, 2), (0.5, -0.5, 0.7))]
    saw_limit = saw_stop = saw_evict = False
    for step, (cam, fwd) in enumerate(path):
        f = O.pool_forward(fwd)
        cands = _ranked_candidates(pool, cam, f)
        victims = _ranked_victims(pool, cam, f)
        before_chunk = pool.slot_chunk.copy()
        n_cand, n, n_ev, loads = pool.update(cam, fwd)
        assert n_cand == len(cands)
        # the pairs: candidate i with victim i while allowed
        want = 0
        while want < min(len(cands), p.max_loads, p.capacity):
            occupied, vi, _ = victims[want]
            if occupied and not vi < -cands[want][0]:
                break
            want += 1
        assert n == want, (step, n, want)
        saw_limit |= n == p.max_loads and len(cands) > p.max_loads
        saw_stop |= n < min(len(cands), p.max_loads)
        nc = pool.nc
        assert [l[0] + nc * (l[1] + nc * l[2]) for l in loads] == [c[1] for c in cands[:n]]
        assert [l[3] for l in loads] == [v[2] for v in victims[:n]]
        assert n_ev == sum(1 for v in victims[:n] if v[0] == 1)
        saw_evict |= n_ev > 0
        # state: inverse maps, stamps, non-empty residents, meshes
        for sl, c in enumerate(pool.slot_chunk):
            if c != O.NO_SLOT:
                assert pool.lookup[c] == sl
        res = np.nonzero(pool.lookup != O.NO_SLOT)[0]
        assert all(pool.slot_chunk[pool.lookup[c]] == c for c in res)
        for x, y, z, sl in loads:
            assert pool.slot_stamp[sl] == step + 1 and before_chunk[sl] != pool.slot_chunk[sl]
            assert O.chunk_nonempty(sc, x, y, z)
            ref = O.mesh_chunk(O.scene_chunk(sc, x, y, z), greedy)
            assert pool.slot_count[sl] == len(ref)
            assert np.array_equal(pool.slot_quads(sl), ref[:p.quads_per_slot])
    assert saw_limit and saw_stop and saw_evict
