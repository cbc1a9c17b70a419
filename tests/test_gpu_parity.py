"""GPU parity through the C-ABI against the CPU oracle on the same seeded inputs.

Integer outputs (hit index, face, quads, voxels, slots, masks) must be equal.
t and shade are compared bit for bit (the tolerance the path promises is 1e-5
relative against FP64; bit equality with the FP32 oracle is stronger), and the
FP64 entry-plane t of every hit is checked to within 1e-5 relative.
"""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def _cam_both(cp, w, h):
    return (R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h),
            O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], w, h))


def _assert_rays_equal(gpu, ref, what):
    idx, face, t, shade = gpu
    ridx, rface, rt, rshade = ref[:4]
    bad = np.nonzero(idx != ridx)[0]
    assert bad.size == 0, "%s: %d hit indices differ, first pixel %d: %d vs %d" % (
        what, bad.size, bad[0], idx[bad[0]], ridx[bad[0]])
    assert np.array_equal(face, rface), what + ": faces differ"
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32)), what + ": t differs"
    assert np.array_equal(shade.view(np.uint32), rshade.view(np.uint32)), what + ": shade differs"


def _fp64_plane_t(cam_o, idx, face, t, n):
    """FP64 t of the entry plane of the chosen voxel face, for every hit."""
    hit = idx >= 0
    cells = np.stack([idx % n, (idx // n) % n, idx // (n * n)], 1)[hit]
    f = face[hit].astype(np.int64)
    axis = np.where(f < 2, 2, np.where(f < 4, 1, 0))
    pos = (f % 2) == 0
    plane = cells[np.arange(len(f)), axis] + pos
    return hit, axis, plane


def test_spheres_dense_voxelization(ctx):
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, W.SPHERES["seed"], 8)
    grid = sc.arrays()["dense"]
    ref = O.spheres_dense(W.SPHERES["seed"], 128, 8).ravel()
    assert np.array_equal(grid, ref)
    assert sc.info().solid_voxels == int((ref > 0).sum())
    sc.destroy()


@pytest.mark.parametrize("flags", [0, O.RAY_SHADOW])
def test_spheres_raymarch_bitexact(ctx, flags):
    cfg = W.SPHERES
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, cfg["seed"], 8)
    cp = W.spheres_camera()
    cam, ocam = _cam_both(cp, 256, 256)
    tg = R.RayTargets(ctx, 256, 256, counters=True)
    R.raymarch(ctx, sc, cam, flags | L.RAY_COUNT, R.shading(W.LIGHT), tg)
    ctx.sync()
    gpu = tg.results()
    osc = O.dense_scene(O.spheres_dense(cfg["seed"], 128, 8))
    ref = O.raymarch(osc, ocam, flags, O.shading(W.LIGHT), threads=8)
    _assert_rays_equal(gpu, ref, "128^3 sphere grid")
    cnt = tg.counts()
    rc = ref[4].as_dict()
    for k in ("n_vox", "n_iter", "hits", "rays", "shadow_rays"):
        assert cnt[k] == rc[k], (k, cnt[k], rc[k])
    sc.destroy()


def test_heightfield_brickmap_voxelization(ctx):
    cfg = W.TERRAIN
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, cfg["n"], cfg["seed"])
    a = sc.arrays()
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, cfg["n"], cfg["seed"])
    assert np.array_equal(a["slots"], slots)
    assert np.array_equal(a["masks"], masks)
    assert np.array_equal(a["voxels"], vox)
    sc.destroy()


def test_heightfield_raymarch_full_frame_bitexact(ctx):
    cfg = W.TERRAIN
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, cfg["n"], cfg["seed"])
    cp = W.terrain_camera()
    w, h = cfg["width"], cfg["height"]
    cam, ocam = _cam_both(cp, w, h)
    tg = R.RayTargets(ctx, w, h, counters=True)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW | L.RAY_COUNT, R.shading(W.LIGHT), tg)
    ctx.sync()
    gpu = tg.results()
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, cfg["n"], cfg["seed"])
    osc = O.brickmap_scene(cfg["n"], slots, masks, vox)
    ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading(W.LIGHT), threads=16)
    _assert_rays_equal(gpu, ref, "512^3 heightfield brick map")
    cnt = tg.counts()
    rc = ref[4].as_dict()
    for k in ("n_brickslot", "n_brickmask", "n_coarse", "n_iter", "hits", "rays", "shadow_rays"):
        assert cnt[k] == rc[k], (k, cnt[k], rc[k])
    # FP64 entry-plane t of the chosen voxel within 1e-5 relative
    idx, face, t, _ = gpu
    hit, axis, plane = _fp64_plane_t(None, idx, face, t, cfg["n"])
    o = np.array(cp["eye"], np.float64)
    px = np.arange(w * h) % w
    py = np.arange(w * h) // w
    sample = np.nonzero(hit)[0][:: max(1, int(hit.sum()) // 20000)]
    for k in sample[:2000]:
        oo, dd = O.primary_ray(ocam, int(px[k]), int(py[k]))
        j = np.searchsorted(np.nonzero(hit)[0], k)
        a = axis[j]
        t64 = (float(plane[j]) - float(oo[a])) / float(dd[a])
        assert abs(float(t[k]) - t64) <= 1e-5 * abs(t64), (k, t[k], t64)
    sc.destroy()


def test_chunk_generation_bitexact(ctx):
    n = 64
    buf = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, W.CHUNKS["seed"], n, buf.ptr, None))
    got = buf.download(np.uint16).reshape(n, 32768)
    ref, _ = O.gen_chunks(W.CHUNKS["seed"], n)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("mode", [L.MESH_CULLED, L.MESH_GREEDY])
def test_mesh_plain_bitexact(ctx, mode):
    n = 16
    buf = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 11, n, buf.ptr, None))
    ref_vox, _ = O.gen_chunks(11, n)
    cnt = ctx.buffer(4)
    q = ctx.buffer(4 * 98304)
    for c in range(n):
        R.mesh_chunk(ctx, buf.ptr + c * 65536, 2, mode, cnt.ptr, q.ptr, 98304)
        ctx.sync()
        k = int(cnt.download(np.uint32)[0])
        ref = O.mesh_chunk(ref_vox[c], mode == L.MESH_GREEDY)
        assert k == len(ref), (c, k, len(ref))
        assert np.array_equal(q.download(np.uint32, k), ref), c


@pytest.mark.parametrize("mode", [L.MESH_CULLED, L.MESH_GREEDY])
def test_mesh_batched_matches_oracle(ctx, mode):
    n = 256
    buf = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 12, n, buf.ptr, None))
    ref_vox, _ = O.gen_chunks(12, n)
    ref_counts, ref_total = O.mesh_chunks_counts(ref_vox, mode == L.MESH_GREEDY, threads=16)
    mt = R.MeshTargets(ctx, n, int(ref_total) + 16)
    R.mesh_chunks(ctx, R.chunk_source(buf, 2, n), mode, mt)
    ctx.sync()
    counts, prefix, total, quads = mt.results()
    assert np.array_equal(counts, ref_counts)
    assert total == ref_total
    assert np.array_equal(prefix, np.concatenate([[0], np.cumsum(ref_counts)[:-1]]).astype(np.uint32))
    for c in range(0, n, 17):
        ref = O.mesh_chunk(ref_vox[c], mode == L.MESH_GREEDY)
        assert np.array_equal(quads[prefix[c]:prefix[c] + counts[c]], ref), c


def test_two_level_caves_generation_and_raymarch_with_ao(ctx):
    n, seed = 256, 5
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, n, seed)
    a = sc.arrays()
    slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, seed)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    assert np.array_equal(a["top"], top)
    assert np.array_equal(a["slots"], ns)
    assert np.array_equal(a["masks"], m2) and np.array_equal(a["voxels"], v2)
    cp = W.caves_camera(n, 320, 180)
    cam, ocam = _cam_both(cp, 320, 180)
    sh = R.shading(W.LIGHT, ao_seed=11, ao_samples=4, ao_radius=16.0)
    tg = R.RayTargets(ctx, 320, 180, index_bytes=8, counters=True)
    flags = L.RAY_SHADOW | L.RAY_AO
    R.raymarch(ctx, sc, cam, flags | L.RAY_COUNT, sh, tg)
    ctx.sync()
    osc = O.two_level_scene(n, top, ns, m2, v2)
    ref = O.raymarch(osc, ocam, flags, O.shading(W.LIGHT, 11, 4, 16.0), threads=8)
    _assert_rays_equal(tg.results(), ref, "two-level caves + AO")
    cnt, rc = tg.counts(), ref[4].as_dict()
    for k in ("n_top", "n_node", "n_brickmask", "hits", "rays", "shadow_rays", "ao_rays", "n_iter"):
        assert cnt[k] == rc[k], (k, cnt[k], rc[k])
    assert (tg.results()[0] >= 0).mean() > 0.3
    sc.destroy()


def test_brickmap_from_host_arrays_odd_edge(ctx):
    # edge 200 is not a multiple of 32 or 128: partial coarse blocks
    n = 200
    slots, masks, vox = O.build_brickmap(O.GEN_TERRAIN_CAVES, n, 77)
    sc = ctx.scene_brickmap(n, slots, masks, vox)
    cp = dict(eye=[-50.0, 180.0, -40.0], target=[100.0, 30.0, 100.0], up=[0.0, 1.0, 0.0], vfov=60.0)
    cam, ocam = _cam_both(cp, 257, 143)
    tg = R.RayTargets(ctx, 257, 143, counters=True)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW | L.RAY_COUNT, R.shading(W.LIGHT), tg)
    ctx.sync()
    ref = O.raymarch(O.brickmap_scene(n, slots, masks, vox), ocam, O.RAY_SHADOW,
                     O.shading(W.LIGHT), threads=8)
    _assert_rays_equal(tg.results(), ref, "host-array brick map, edge 200")
    cnt, rc = tg.counts(), ref[4].as_dict()
    for k in ("n_brickslot", "n_brickmask", "n_coarse", "n_iter", "hits", "shadow_rays"):
        assert cnt[k] == rc[k], (k, cnt[k], rc[k])
    sc.destroy()


def test_two_level_from_host_arrays(ctx):
    n = 256
    g = np.zeros((n, n, n), np.uint16)
    g[10:60, 140:200, 30:90] = 3
    g[200:230, 5:40, 150:250] = 4
    slots, masks, vox = O.brickmap_from_dense(g)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    sc = ctx.scene_two_level(n, top, ns, m2, v2)
    cp = dict(eye=[-100.0, 300.0, -80.0], target=[128.0, 100.0, 128.0], up=[0.0, 1.0, 0.0], vfov=50.0)
    cam, ocam = _cam_both(cp, 200, 150)
    tg = R.RayTargets(ctx, 200, 150, index_bytes=8)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    ctx.sync()
    ref = O.raymarch(O.two_level_scene(n, top, ns, m2, v2), ocam, O.RAY_SHADOW, O.shading(W.LIGHT))
    _assert_rays_equal(tg.results(), ref, "host-array two-level")
    sc.destroy()
