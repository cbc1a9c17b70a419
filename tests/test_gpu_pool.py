"""The chunk pool on the GPU against the oracle's pool: after every update
of a camera path, the counts, the lookup table, the slots (chunk, stamp,
quad count, quads) and the load list are bit-identical; edits are remeshed
into the resident slots; argument checks."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R

pytestmark = pytest.mark.gpu

PATH = [((1, 1, 1), (1, 0, 0)), ((1, 1, 1), (1, 0, 0)), ((2, 1, 1), (1, 0.2, 0)),
        ((3, 2, 2), (0, 0, 1)), ((5, 2, 5), (-1, 0, -0.5)), ((6, 1, 6), (0, 1, 0)),
        ((6, 1, 6), (0, 1, 0)), ((2, 2, 2), (0.5, -0.5, 0.7)), ((-3, 9, 4), (1, -1, 0))]


def _oracle_scene(kind, gen, n, seed):
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    if kind == L.SCENE_TWO_LEVEL_U16:
        top, slots, masks, vox = O.two_level_from_bricks(n, slots, masks, vox)
    masks, vox = np.ascontiguousarray(masks.copy()), np.ascontiguousarray(vox.copy())
    if kind == L.SCENE_BRICKMAP_U16:
        return O.brickmap_scene(n, slots, masks, vox), masks, vox
    return O.two_level_scene(n, top, slots, masks, vox), masks, vox


def _compare(pool, opool, step):
    st = pool.stats()
    g = pool.state()
    assert st["stamp"] == opool.stamp, step
    assert np.array_equal(g["lookup"], opool.lookup), step
    assert np.array_equal(g["slot_chunk"], opool.slot_chunk), step
    assert np.array_equal(g["slot_stamp"], opool.slot_stamp), step
    res = opool.slot_chunk != O.NO_SLOT
    assert np.array_equal(g["slot_count"][res], opool.slot_count[res]), step
    q = opool.p.quads_per_slot
    for sl in np.nonzero(res)[0]:
        k = min(int(opool.slot_count[sl]), q)
        assert np.array_equal(g["quads"][sl * q: sl * q + k], opool.quads[sl * q: sl * q + k]), \
            (step, sl)
    assert st["n_resident"] == int(res.sum())
    assert st["n_quads"] == int(opool.slot_count[res].astype(np.uint64).sum())
    assert st["n_truncated"] == int((opool.slot_count[res] > q).sum())
    return st, g


@pytest.mark.parametrize("kind,gen,n,seed,mode,bake", [
    (L.SCENE_BRICKMAP_U16, L.GEN_HEIGHTFIELD, 256, 2, L.MESH_GREEDY, 0),
    (L.SCENE_BRICKMAP_U16, L.GEN_HEIGHTFIELD, 256, 2, L.MESH_CULLED, 0),
    (L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, 256, 5, L.MESH_GREEDY, 0),
    (L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, 256, 5, L.MESH_GREEDY, 256)])
def test_pool_updates_match_oracle(ctx, kind, gen, n, seed, mode, bake):
    sc = ctx.scene_generate(gen, kind, n, seed)
    osc, _, _ = _oracle_scene(kind, gen, n, seed)
    kw = dict(capacity=48, quads_per_slot=3000, forward_radius=4, backward_radius=2,
              view_angle=100.0, max_loads=12)
    pool = R.ChunkPool(ctx, sc, mode=mode, bake_directions=bake, **kw)
    opool = O.OraclePool(osc, O.pool_params(greedy=mode == L.MESH_GREEDY, **kw),
                         bake_directions=bake)
    loads_seen = evicted_seen = truncated_seen = 0
    for step, (cam, fwd) in enumerate(PATH):
        pool.update(cam, fwd)
        n_cand, nl, nev, loads = opool.update(cam, fwd)
        st, g = _compare(pool, opool, step)
        assert (st["n_candidates"], st["n_loads"], st["n_evicted"]) == (n_cand, nl, nev), step
        for i, (x, y, z, sl) in enumerate(loads):
            assert list(g["load_xyz"][3 * i:3 * i + 3]) == [x, y, z] and g["load_slot"][i] == sl
        loads_seen += nl
        evicted_seen += nev
        truncated_seen = max(truncated_seen, st["n_truncated"])
    assert loads_seen > 48 and evicted_seen > 0
    pool.destroy()
    sc.destroy()


def test_pool_remesh_after_edits(ctx):
    n, seed = 256, 5
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, seed)
    osc, _, _ = _oracle_scene(L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, n, seed)
    kw = dict(capacity=200, quads_per_slot=2048, forward_radius=5, backward_radius=5,
              view_angle=120.0, max_loads=200)
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **kw)
    opool = O.OraclePool(osc, O.pool_params(greedy=True, **kw))
    pool.update((3, 3, 3), (0, 1, 0))
    opool.update((3, 3, 3), (0, 1, 0))
    _compare(pool, opool, "load")
    # edit resident and non-resident chunks and one outside, then repeat two
    # of them in a second call (the chunks of one call are distinct)
    nc = opool.nc
    res = [int(c) for c in opool.slot_chunk if c != O.NO_SLOT]
    xyz = [[c % nc, (c // nc) % nc, c // (nc * nc)] for c in res[::3]]
    xyz += [[0, 0, 7], [7, 7, 0], [-1, 0, 0]]
    xyz = np.asarray(xyz, np.int32)
    again = np.ascontiguousarray(xyz[[0, 1]])
    d, d2 = ctx.upload(xyz.ravel()), ctx.upload(again.ravel())
    L.check(ctx.lib.meso_scene_edit_chunks(sc.h, d.ptr, len(xyz), 77, None))
    L.check(ctx.lib.meso_scene_edit_chunks(sc.h, d2.ptr, len(again), 91, None))
    O.edit_chunks(osc, xyz, 77)
    O.edit_chunks(osc, again, 91)
    pool.remesh(d, len(xyz))
    ctx.sync()
    q = kw["quads_per_slot"]
    g = pool.state()
    changed = 0
    for c in res:
        sl = int(opool.lookup[c])
        ref = O.mesh_chunk(O.scene_chunk(osc, c % nc, (c // nc) % nc, c // (nc * nc)), True)
        assert g["slot_count"][sl] == len(ref), c
        k = min(len(ref), q)
        assert np.array_equal(g["quads"][sl * q: sl * q + k], ref[:k]), c
        changed += int(opool.slot_count[sl]) != len(ref)
    assert changed > 0, "the edits changed no resident mesh"
    pool.destroy()
    sc.destroy()


def test_pool_argument_checks(ctx):
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, 128, 2)
    h = C.c_void_p()
    good = dict(capacity=64, quads_per_slot=512, forward_radius=4, backward_radius=2,
                view_angle=90.0, max_loads=16, mesh_mode=L.MESH_GREEDY)
    for bad in [dict(capacity=0), dict(quads_per_slot=0), dict(forward_radius=0),
                dict(backward_radius=5), dict(view_angle=0.0), dict(view_angle=float("nan")),
                dict(max_loads=65), dict(mesh_mode=7), dict(bake_directions=1),
                dict(bake_directions=70000)]:
        kw = dict(good, **bad)
        d = L.PoolDesc(C.sizeof(L.PoolDesc), **kw)
        assert ctx.lib.meso_chunk_pool_create(ctx.h, sc.h, L.byref(d), C.byref(h)) == L.ERR_ARG, bad
    dense = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    d = L.PoolDesc(C.sizeof(L.PoolDesc), **good)
    assert ctx.lib.meso_chunk_pool_create(ctx.h, dense.h, L.byref(d), C.byref(h)) == \
        L.ERR_NOT_SUPPORTED
    pool = R.ChunkPool(ctx, sc, **{k: v for k, v in good.items() if k != "mesh_mode"})
    cam = (C.c_int32 * 3)(0, 0, 0)
    for fwd in [(0.0, 0.0, 0.0), (float("nan"), 1.0, 0.0), (float("inf"), 0.0, 0.0)]:
        f = (C.c_float * 3)(*fwd)
        assert ctx.lib.meso_chunk_pool_update(pool.h, C.cast(cam, C.c_void_p),
                                              C.cast(f, C.c_void_p), None) == L.ERR_ARG
    far = (C.c_int32 * 3)(1 << 25, 0, 0)
    f = (C.c_float * 3)(1.0, 0.0, 0.0)
    assert ctx.lib.meso_chunk_pool_update(pool.h, C.cast(far, C.c_void_p), C.cast(f, C.c_void_p),
                                          None) == L.ERR_ARG
    # a camera far outside the scene loads nothing
    pool.update((100, 100, 100), (1, 0, 0))
    assert pool.stats()["n_loads"] == 0
    pool.destroy()
    dense.destroy()
    sc.destroy()


@pytest.mark.parametrize("capacity,radius,max_loads", [
    (20000, 4, 64),     # more slots than the selection holds in registers
    (5000, 9, 4096),    # the selection's largest limit: 4,096 places sorted in LDS
    (5000, 9, 5000),    # a limit above the selection's, so the lists are sorted
    (140000, 4, 64),    # more slots than the selection takes: the lists are sorted
    (5000, 26, 64)])    # a cube of 53^3 offsets, longer than the selection takes
def test_pool_selection_paths(ctx, capacity, radius, max_loads):
    n, seed = 256, 5
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, seed)
    osc, _, _ = _oracle_scene(L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, n, seed)
    kw = dict(capacity=capacity, quads_per_slot=256, forward_radius=radius, backward_radius=2,
              view_angle=100.0, max_loads=max_loads)
    pool = R.ChunkPool(ctx, sc, mode=L.MESH_GREEDY, **kw)
    opool = O.OraclePool(osc, O.pool_params(greedy=True, **kw))
    for step, (cam, fwd) in enumerate(PATH[:5]):
        pool.update(cam, fwd)
        n_cand, nl, nev, loads = opool.update(cam, fwd)
        st, g = _compare(pool, opool, step)
        assert (st["n_candidates"], st["n_loads"], st["n_evicted"]) == (n_cand, nl, nev), step
        for i, (x, y, z, sl) in enumerate(loads):
            assert list(g["load_xyz"][3 * i:3 * i + 3]) == [x, y, z] and g["load_slot"][i] == sl
    pool.destroy()
    sc.destroy()
