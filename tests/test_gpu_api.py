"""GPU tests of the rest of the C-ABI: dense-grid chunk meshing, graph replay,
the autotuner, the reference sphere voxelizer with its erosion chain, brick
edits and argument errors."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def tiles32(grid):
    """[z][y][x] dense grid -> chunks in linear chunk order (cx fastest)"""
    n = grid.shape[0]
    nc = n // 32
    out = []
    for cz in range(nc):
        for cy in range(nc):
            for cx in range(nc):
                out.append(np.ascontiguousarray(
                    grid[cz * 32:(cz + 1) * 32, cy * 32:(cy + 1) * 32, cx * 32:(cx + 1) * 32]).ravel())
    return out


@pytest.mark.parametrize("mode", [L.MESH_CULLED, L.MESH_GREEDY])
def test_dense_grid_u8_chunks(ctx, mode):
    # the 128^3 u8 sphere grid meshed as 4^3 chunks of 32^3
    g = O.spheres_dense(W.SPHERES["seed"], 128, 8)
    buf = ctx.upload(g.ravel())
    chunks = tiles32(g)
    refs = [O.mesh_chunk(c, mode == L.MESH_GREEDY) for c in chunks]
    total = sum(len(r) for r in refs)
    mt = R.MeshTargets(ctx, 64, total)
    R.mesh_chunks(ctx, R.chunk_source(buf, 1, 64, L.CHUNKS_DENSE_GRID, 128), mode, mt)
    ctx.sync()
    counts, prefix, tot, quads = mt.results()
    assert tot == total
    for c in range(64):
        assert np.array_equal(quads[prefix[c]:prefix[c] + counts[c]], refs[c]), c


@pytest.mark.parametrize("mode", [L.MESH_CULLED, L.MESH_GREEDY])
def test_batched_mesh_past_capacity_keeps_the_prefix(ctx, mode):
    """A quad buffer smaller than the batch's quads: the counts, prefixes and
    total are those of the whole batch, and the buffer holds the first
    `capacity` quads of the whole output (planes that straddle the end are
    written up to it, later ones not at all)."""
    n = 64
    vox = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 31, n, vox.ptr, None))
    src = R.chunk_source(vox, 2, n)
    full = R.MeshTargets(ctx, n, 4 << 20)
    R.mesh_chunks(ctx, src, mode, full)
    ctx.sync()
    fc, fp, ft, fq = full.results()
    assert 1000 < ft <= 4 << 20
    ref_vox, _ = O.gen_chunks(31, n)
    for c in range(0, n, 9):
        assert np.array_equal(fq[fp[c]:fp[c] + fc[c]], O.mesh_chunk(ref_vox[c], mode == L.MESH_GREEDY))
    for cap in (ft // 3 + 7, ft - 1):
        mt = R.MeshTargets(ctx, n, cap)
        R.mesh_chunks(ctx, src, mode, mt)
        ctx.sync()
        c, p, t, q = mt.results()
        assert t == ft and np.array_equal(c, fc) and np.array_equal(p, fp), cap
        assert np.array_equal(q, fq[:cap]), cap


def _wide_u16_chunks():
    """u16 chunks whose materials use all 16 bits: neighbours that differ only
    in the high byte, 0x8000 and 0xFFFF, runs ending on the 8-voxel piece
    boundaries and at the chunk's last row and slice."""
    rng = np.random.default_rng(77)
    pool = np.array([0, 1, 0x00FF, 0x0100, 0x0101, 0x7FFF, 0x8000, 0x8001, 0xFFFF], np.uint16)
    out = []
    a = pool[rng.integers(0, len(pool), 32768)]  # independent voxels
    out.append(a)
    b = np.repeat(pool[rng.integers(0, len(pool), 32768 // 4)], 4)  # runs of 4 along x
    out.append(b)
    c = np.zeros((32, 32, 32), np.uint16)  # (z, y, x): stripes differing in the high byte only
    c[:, :, 0::2] = 0x0001
    c[:, :, 1::2] = 0x0101
    c[::3, :, :] = 0x8001
    c[:, 31, :] = 0xFFFF
    out.append(c.ravel())
    d = np.full((32, 32, 32), 0xFFFF, np.uint16)  # solid, one material, holes at the edges
    d[31, :, 7::8] = 0
    d[:, 31, 8::8] = 0x7FFF
    out.append(d.ravel())
    e = np.zeros((32, 32, 32), np.uint16)  # blocks whose material varies along y and z only
    for z in range(32):
        for y in range(32):
            e[z, y, (y + z) % 8:(y + z) % 8 + 20] = 0x0100 * ((y // 4) % 2) + 0x0001 * ((z // 8) % 2) + 1
    out.append(e.ravel())
    f = pool[rng.integers(0, len(pool), 32768)] * (rng.random(32768) < 0.15)  # sparse
    out.append(f.astype(np.uint16))
    return np.stack(out)


@pytest.mark.parametrize("mode", [L.MESH_CULLED, L.MESH_GREEDY])
def test_wide_u16_materials_match_oracle(ctx, mode):
    """The row masks of u16 chunks compare whole 16-bit materials: the batched
    and the per-chunk path against the oracle on materials that differ only in
    their high byte or use the top bit."""
    chunks = _wide_u16_chunks()
    n = chunks.shape[0]
    refs = [O.mesh_chunk(chunks[i], mode == L.MESH_GREEDY) for i in range(n)]
    total = sum(len(r) for r in refs)
    buf = ctx.upload(chunks.ravel())
    mt = R.MeshTargets(ctx, n, total)
    R.mesh_chunks(ctx, R.chunk_source(buf, 2, n), mode, mt)
    ctx.sync()
    counts, prefix, tot, quads = mt.results()
    assert tot == total
    cnt = ctx.buffer(4)
    q = ctx.buffer(4 * 98304)
    for i in range(n):
        assert counts[i] == len(refs[i]), i
        assert np.array_equal(quads[prefix[i]:prefix[i] + counts[i]], refs[i]), i
        R.mesh_chunk(ctx, buf.ptr + i * 65536, 2, mode, cnt.ptr, q.ptr, 98304)
        ctx.sync()
        k = int(cnt.download(np.uint32)[0])
        assert k == len(refs[i]) and np.array_equal(q.download(np.uint32, k), refs[i]), i


def test_graph_replay_matches_eager(ctx):
    n = 128
    vox = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 31, n, vox.ptr, None))
    src = R.chunk_source(vox, 2, n)
    mt = R.MeshTargets(ctx, n, 4 << 20)
    R.mesh_chunks(ctx, src, L.MESH_GREEDY, mt)
    ctx.sync()
    eager = mt.results()
    g = C.c_void_p()
    gt = R.MeshTargets(ctx, n, 4 << 20)
    L.check(ctx.lib.meso_mesh_graph_create(ctx.h, L.byref(src), L.MESH_GREEDY, gt.counts.ptr,
                                           gt.prefix.ptr, gt.quads.ptr, gt.capacity, gt.total.ptr,
                                           gt.work.ptr, C.byref(g)))
    for rep in range(3):
        gt.counts.zero()
        gt.quads.zero()
        L.check(ctx.lib.meso_graph_launch(g, None))
        ctx.sync()
        got = gt.results()
        for a, b in zip(eager, got):
            assert np.array_equal(np.asarray(a), np.asarray(b)), rep
        # eager work between replays
        R.mesh_chunks(ctx, src, L.MESH_CULLED, mt)
        ctx.sync()
    L.check(ctx.lib.meso_graph_destroy(g))
    # plain graph: one launch node per chunk
    per = int(eager[0].max())
    pc, pq = ctx.buffer(4 * n), ctx.buffer(4 * per * n)
    L.check(ctx.lib.meso_mesh_plain_graph_create(ctx.h, L.byref(src), L.MESH_GREEDY, pc.ptr, pq.ptr,
                                                 per, C.byref(g)))
    for rep in range(2):
        L.check(ctx.lib.meso_graph_launch(g, None))
        ctx.sync()
        assert np.array_equal(pc.download(np.uint32), eager[0])
    q = pq.download(np.uint32).reshape(n, per)
    for c in range(0, n, 9):
        k = int(eager[0][c])
        assert np.array_equal(q[c, :k], eager[3][eager[1][c]:eager[1][c] + k])
    L.check(ctx.lib.meso_graph_destroy(g))


def test_autotuners_keep_results(ctx, tmp_path):
    n = 64
    vox = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 32, n, vox.ptr, None))
    src = R.chunk_source(vox, 2, n)
    mt = R.MeshTargets(ctx, n, 2 << 20)
    R.mesh_chunks(ctx, src, L.MESH_GREEDY, mt)
    ctx.sync()
    before = mt.results()
    tr = L.TuneResult()
    tr.size = C.sizeof(tr)
    L.check(ctx.lib.meso_autotune_mesh(ctx.h, L.byref(src), L.MESH_GREEDY, mt.counts.ptr,
                                       mt.prefix.ptr, mt.quads.ptr, mt.capacity, mt.total.ptr,
                                       mt.work.ptr, L.byref(tr)))
    assert tr.tried >= 5 and tr.best_ms > 0 and tr.best_ms <= tr.default_ms * 1.01
    R.mesh_chunks(ctx, src, L.MESH_GREEDY, mt)  # now with the tuned shape
    ctx.sync()
    for a, b in zip(before, mt.results()):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    cp = W.spheres_camera(width=200, height=120)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 200, 120)
    tg = R.RayTargets(ctx, 200, 120)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(), tg)
    ctx.sync()
    ref = tg.results()
    rr = L.TuneResult()
    rr.size = C.sizeof(rr)
    L.check(ctx.lib.meso_autotune_raymarch(ctx.h, sc.h, L.byref(cam), L.RAY_SHADOW,
                                           L.byref(R.shading()), L.byref(tg.out), L.byref(rr)))
    assert rr.tried == 14 and rr.tile_w * rr.tile_h == rr.block
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(), tg)
    ctx.sync()
    for a, b in zip(ref, tg.results()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    path = str(tmp_path / "tune.txt").encode()
    L.check(ctx.lib.meso_autotune_save(ctx.h, path))
    assert os.path.getsize(path) > 0
    L.check(ctx.lib.meso_autotune_load(ctx.h, path))
    sc.destroy()


def test_persistent_shapes_write_every_pixel_after_the_tuner():
    """A fresh context (its band counters allocated where a freed buffer held
    other bytes) tunes a frame, then renders it in every persistent shape into
    poisoned outputs: every pixel is written and equals the oracle's."""
    c2 = R.Context(0)
    try:
        _render_after_tuner(c2)  # its buffers are freed on return, before the context
    finally:
        c2.close()


def _render_after_tuner(c2):
    junk = c2.buffer(16 << 20, np.full(16 << 20, 0xFF, np.uint8))
    del junk
    sc = c2.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    cp = W.spheres_camera(width=200, height=120)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 200, 120)
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 200, 120)
    ref = O.raymarch(O.dense_scene(sc.arrays()["dense"]), ocam, O.RAY_SHADOW, O.shading())
    tg = R.RayTargets(c2, 200, 120)
    rr = L.TuneResult()
    rr.size = C.sizeof(rr)
    L.check(c2.lib.meso_autotune_raymarch(c2.h, sc.h, L.byref(cam), L.RAY_SHADOW,
                                          L.byref(R.shading()), L.byref(tg.out), L.byref(rr)))
    cus = c2.device_info().compute_units
    tiles = {5: (8, 8), 6: (16, 4), 7: (32, 2), 8: (8, 8), 9: (16, 4)}
    for v, g in ((5, 0), (5, 16 * cus), (6, 0), (7, 0), (8, 0), (8, 16 * cus), (9, 0)):
        pin = L.TuneResult()
        pin.size = C.sizeof(pin)
        pin.variant, pin.grid = v, g
        pin.tile_w, pin.tile_h = tiles[v]
        L.check(c2.lib.meso_autotune_pin_raymarch(c2.h, sc.h, L.byref(cam), L.RAY_SHADOW,
                                                  L.byref(pin)))
        for b in (tg.index, tg.face, tg.t, tg.shade):
            b.upload(np.full(b.nbytes, 0xA5, np.uint8))
        R.raymarch(c2, sc, cam, L.RAY_SHADOW, R.shading(), tg)
        c2.sync()
        for a, b in zip(tg.results(), ref[:4]):
            b = np.ascontiguousarray(np.asarray(b, a.dtype))
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint8), b.view(np.uint8)), (v, g)
    sc.destroy()


def test_reference_sphere_chunks_and_erosion(ctx):
    coords = np.array([(cx, cy, cz) for cz in range(-4, 4) for cy in range(-4, 4)
                       for cx in range(3, 10)], np.int32)
    nc = len(coords)
    dc = ctx.upload(coords.ravel())
    mips = ctx.buffer(nc * 4 * 4096)
    L.check(ctx.lib.meso_generate_sphere_chunks(ctx.h, dc.ptr, nc, 16, 1.0, 4, mips.ptr, None))
    ctx.sync()
    got = mips.download(np.uint8).reshape(nc, 4, 16, 16, 16)
    solid = 0
    for i, (cx, cy, cz) in enumerate(coords):
        cells = O.ref_generate_sphere(int(cx), int(cy), int(cz), 1.0, 16)
        occ = np.zeros((16, 16, 16), np.uint8)
        if len(cells):
            occ[cells[:, 2], cells[:, 1], cells[:, 0]] = 1
        solid += len(cells)
        assert np.array_equal(got[i], O.ref_erode_mips(occ, 4)), (cx, cy, cz)
    xs = np.arange(48, 160)[:, None, None]
    ys = np.arange(-64, 64)[None, :, None]
    zs = np.arange(-64, 64)[None, None, :]
    assert solid == int(((xs - 100) ** 2 + ys ** 2 + zs ** 2 < 2500).sum())


def test_update_bricks(ctx):
    n = 128
    sc = ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, n, 60)
    a = sc.arrays()
    nb = len(a["masks"])
    rng = np.random.default_rng(5)
    sel = rng.choice(nb, 20, replace=False).astype(np.uint32)
    edits = rng.integers(0, 4, (20, 512)).astype(np.uint16) * (rng.random((20, 512)) < 0.5)
    edits[0] = 0  # a brick emptied by the edit keeps its slot with a zero mask
    ds, dv = ctx.upload(sel), ctx.upload(edits.astype(np.uint16).ravel())
    L.check(ctx.lib.meso_scene_update_bricks(sc.h, ds.ptr, dv.ptr, 20, None))
    b = sc.arrays()
    vox = a["voxels"].copy()
    vox[sel] = edits
    masks = a["masks"].copy()
    for i, s in enumerate(sel):
        v = edits[i].reshape(8, 8, 8)
        for lz in range(8):
            m = 0
            for ly in range(8):
                for lx in range(8):
                    if v[lz, ly, lx]:
                        m |= 1 << (lx + 8 * ly)
            masks[s, lz] = m
    assert np.array_equal(b["voxels"], vox) and np.array_equal(b["masks"], masks)
    assert np.array_equal(b["slots"], a["slots"])
    cp = dict(eye=[-30.0, 120.0, -20.0], target=[64.0, 20.0, 64.0], up=[0.0, 1.0, 0.0], vfov=60.0)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], 60.0, 160, 100)
    tg = R.RayTargets(ctx, 160, 100)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    ctx.sync()
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], 60.0, 160, 100)
    ref = O.raymarch(O.brickmap_scene(n, b["slots"], b["masks"], b["voxels"]), ocam, O.RAY_SHADOW,
                     O.shading(W.LIGHT))
    for x, y in zip(tg.results(), ref[:4]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    bad = ctx.upload(np.array([nb + 5], np.uint32))
    with pytest.raises(L.MesoError) as e:
        L.check(ctx.lib.meso_scene_update_bricks(sc.h, bad.ptr, dv.ptr, 1, None))
    assert e.value.status == L.ERR_ARG
    sc.destroy()


def test_host_bricks_unreferenced_and_held_twice(ctx):
    """A brick no cell holds can be updated (it stays invisible); a brick
    named by two cells is refused at creation."""
    n = 128
    g = ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, n, 60)
    a = g.arrays()
    g.destroy()
    nb = len(a["masks"])
    rng = np.random.default_rng(9)
    spare_v = (rng.integers(1, 4, 512) * (rng.random(512) < 0.5)).astype(np.uint16)
    spare_m = np.zeros(8, np.uint64)
    for i in np.nonzero(spare_v)[0]:
        spare_m[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
    masks = np.concatenate([a["masks"], spare_m[None]])
    vox = np.concatenate([a["voxels"], spare_v[None]])
    sc = ctx.scene_brickmap(n, a["slots"], masks, vox)  # brick nb: held by no cell
    edits = np.stack([np.full(512, 3, np.uint16), np.zeros(512, np.uint16)])
    sel = np.array([nb, int(a["slots"][a["slots"] != 0xFFFFFFFF][7])], np.uint32)
    ds, dv = ctx.upload(sel), ctx.upload(edits.ravel())
    L.check(ctx.lib.meso_scene_update_bricks(sc.h, ds.ptr, dv.ptr, 2, None))
    b = sc.arrays()
    assert np.array_equal(b["voxels"][sel], edits)
    cp = dict(eye=[-30.0, 120.0, -20.0], target=[64.0, 20.0, 64.0], up=[0.0, 1.0, 0.0])
    cam = R.camera(cp["eye"], cp["target"], cp["up"], 60.0, 160, 100)
    tg = R.RayTargets(ctx, 160, 100)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    ctx.sync()
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], 60.0, 160, 100)
    ref = O.raymarch(O.brickmap_scene(n, b["slots"], b["masks"], b["voxels"]), ocam, O.RAY_SHADOW,
                     O.shading(W.LIGHT))
    for x, y in zip(tg.results(), ref[:4]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    sc.destroy()
    twice = a["slots"].copy()
    held = np.nonzero(twice != 0xFFFFFFFF)[0]
    empty = np.nonzero(twice == 0xFFFFFFFF)[0]
    twice[empty[0]] = twice[held[0]]
    h = C.c_void_p()
    m = np.ascontiguousarray(a["masks"])
    v = np.ascontiguousarray(a["voxels"])
    assert ctx.lib.meso_scene_create_brickmap_u16(ctx.h, n, twice.ctypes.data, nb, m.ctypes.data,
                                                  v.ctypes.data, C.byref(h)) == L.ERR_ARG
    assert b"two cells" in ctx.lib.meso_last_error_message()


def test_frames_in_flight_on_two_streams(ctx):
    """Consecutive frames on two streams, ordered by events, each into its own
    outputs, equal the frame rendered alone."""
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, 256, 2)
    cp = W.terrain_camera(256, 320, 200)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 320, 200)
    ref = R.RayTargets(ctx, 320, 200)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), ref)
    ctx.sync()
    want = ref.results()
    s2 = C.c_void_p()
    L.check(ctx.lib.meso_stream_create(ctx.h, C.byref(s2)))
    tgs = [R.RayTargets(ctx, 320, 200), R.RayTargets(ctx, 320, 200)]
    streams = [None, s2.value]
    start, mid = R.Event(ctx), R.Event(ctx)
    start.record()
    L.check(ctx.lib.meso_stream_wait_event(ctx.h, s2, start.h))
    for i in range(6):
        R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tgs[i % 2], streams[i % 2])
    mid.record(s2)
    L.check(ctx.lib.meso_stream_wait_event(ctx.h, None, mid.h))
    ctx.sync()
    for tg in tgs:
        for x, y in zip(tg.results(), want):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    assert ctx.lib.meso_stream_wait_event(ctx.h, s2, None) == L.ERR_ARG
    L.check(ctx.lib.meso_stream_destroy(ctx.h, s2))
    sc.destroy()


def test_argument_errors(ctx):
    lib = ctx.lib
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 64, 3, 4)
    tg = R.RayTargets(ctx, 8, 8)
    cam = R.camera((100, 100, 100), (32, 32, 32), (0, 0, 1), 60.0, 8, 8)
    bad_up = R.camera((0, 0, 100), (0, 0, 0), (0, 0, 1), 60.0, 8, 8)
    assert lib.meso_raymarch(ctx.h, sc.h, L.byref(bad_up), 0, None, L.byref(tg.out), None) == L.ERR_ARG
    nan_eye = R.camera((float("nan"), 100, 100), (32, 32, 32), (0, 0, 1), 60.0, 8, 8)
    assert lib.meso_raymarch(ctx.h, sc.h, L.byref(nan_eye), 0, None, L.byref(tg.out), None) == L.ERR_ARG
    assert b"finite" in lib.meso_last_error_message()
    tg.out.index_bytes = 3
    assert lib.meso_raymarch(ctx.h, sc.h, L.byref(cam), 0, None, L.byref(tg.out), None) == L.ERR_ARG
    tg.out.index_bytes = 4
    sh = R.shading(ao_samples=0)
    assert lib.meso_raymarch(ctx.h, sc.h, L.byref(cam), L.RAY_AO, L.byref(sh), L.byref(tg.out),
                             None) == L.ERR_ARG
    assert lib.meso_raymarch(ctx.h, sc.h, L.byref(cam), L.RAY_COUNT, None, L.byref(tg.out),
                             None) == L.ERR_ARG
    buf = ctx.buffer(2 * 32768 + 64)
    cnt = ctx.buffer(4)
    assert lib.meso_mesh_chunk(ctx.h, buf.ptr + 2, 2, L.MESH_GREEDY, cnt.ptr, None, 0,
                               None) == L.ERR_ARG
    slots = np.full(8 ** 3, 0xFFFFFFFF, np.uint32)
    slots[5] = 3  # names a brick that does not exist
    h = C.c_void_p()
    m = np.zeros((1, 8), np.uint64)
    v = np.zeros((1, 512), np.uint16)
    assert lib.meso_scene_create_brickmap_u16(ctx.h, 64, slots.ctypes.data, 1, m.ctypes.data,
                                              v.ctypes.data, C.byref(h)) == L.ERR_ARG
    assert b"slot" in lib.meso_last_error_message()
    assert lib.meso_buffer_upload(buf.h, m.ctypes.data, 1 << 30, 0) == L.ERR_ARG
    assert lib.meso_buffer_device_address(buf.h, 3) == 0
    # a stream's raymarch counters go with the stream; a new stream renders the same bits
    pin = L.TuneResult()
    pin.size = C.sizeof(pin)
    pin.variant, pin.tile_w, pin.tile_h, pin.grid = 5, 8, 8, 0  # persistent waves
    L.check(lib.meso_autotune_pin_raymarch(ctx.h, sc.h, L.byref(cam), 0, L.byref(pin)))
    for _ in range(2):
        st = C.c_void_p()
        L.check(lib.meso_stream_create(ctx.h, C.byref(st)))
        tg2 = R.RayTargets(ctx, 8, 8)
        L.check(lib.meso_raymarch(ctx.h, sc.h, L.byref(cam), 0, None, L.byref(tg2.out), st))
        L.check(lib.meso_stream_sync(ctx.h, st))
        L.check(lib.meso_stream_destroy(ctx.h, st))
        R.raymarch(ctx, sc, cam, 0, None, tg)
        ctx.sync()
        for a, b in zip(tg.results(), tg2.results()):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    sc.destroy()


def test_pinned_host_copies(ctx):
    data = np.arange(1 << 20, dtype=np.uint32)
    buf = ctx.upload(data)
    p = C.c_void_p()
    L.check(ctx.lib.meso_host_alloc(ctx.h, data.nbytes, C.byref(p)))
    try:
        L.check(ctx.lib.meso_memcpy_d2h(ctx.h, p, buf.ptr, data.nbytes, None))
        got = np.ctypeslib.as_array((C.c_uint32 * data.size).from_address(p.value))
        assert np.array_equal(got, data)
        dst = ctx.buffer(data.nbytes)
        L.check(ctx.lib.meso_memcpy_d2d(ctx.h, dst.ptr, buf.ptr, data.nbytes, None))
        ctx.sync()
        assert np.array_equal(dst.download(np.uint32), data)
    finally:
        L.check(ctx.lib.meso_host_free(ctx.h, p))
    assert ctx.lib.meso_host_alloc(ctx.h, 0, C.byref(p)) == L.ERR_ARG


def test_empty_and_single_brick_scenes(ctx):
    """A brick scene with no brick (its solid box is empty) and one with a
    single brick: the raymarcher, the chunk pool and the quad pass all agree
    with the oracle (which clips to the same box)."""
    for fill in (None, (9, 17, 3)):
        g = np.zeros((64, 64, 64), np.uint16)  # (z, y, x)
        if fill:
            x, y, z = fill
            g[z, y, x] = 5
        slots, masks, vox = O.brickmap_from_dense(g)
        sc = ctx.scene_brickmap(64, slots, masks, vox)
        osc = O.brickmap_scene(64, slots, masks, vox)
        cam = R.camera([14.3, 21.7, -2.1], [9.5, 17.5, 3.5], [0.0, 1.0, 0.0], 40.0, 96, 64)
        tg = R.RayTargets(ctx, 96, 64)
        R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(), tg)
        ctx.sync()
        ocam = O.camera([14.3, 21.7, -2.1], [9.5, 17.5, 3.5], [0.0, 1.0, 0.0], 40.0, 96, 64)
        ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading())
        got = tg.results()
        for a, b in zip(got, ref[:4]):
            assert a.astype(b.dtype).tobytes() == b.tobytes()
        assert ((got[0] >= 0).sum() > 50) == bool(fill)
        pool = R.ChunkPool(ctx, sc, capacity=8, quads_per_slot=64, forward_radius=3,
                           backward_radius=3, view_angle=120.0, max_loads=8)
        pool.update((0, 0, 0), (1, 0, 0))
        assert pool.stats()["n_resident"] == (1 if fill else 0)
        rast = R.Raster(ctx, 96, 64)
        tr = R.RayTargets(ctx, 96, 64)
        rast.draw(pool, cam, R.shading(), tr)
        ctx.sync()
        ri, rf, rt, _ = tr.results()
        same = (ri == got[0]) & (rf == got[1])
        assert same.mean() >= 0.99  # pixels whose rays graze the voxel's edges may differ
        assert np.array_equal(rt.view(np.uint32)[same], got[2].view(np.uint32)[same])
        rast.destroy()
        pool.destroy()
        sc.destroy()
