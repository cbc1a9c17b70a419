"""The HIP path, through the C-ABI, reproduces the committed golden fixtures
(tests/golden/*.npz, written by tests/golden/make_golden.py with the oracle)."""
import hashlib
import os

import numpy as np
import pytest

from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
LIGHT = (0.4, 1.0, 0.3)


def _load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _frame(ctx, sc, g, flags, sh, index_bytes=4, counters=False):
    w, h = (int(v) for v in g["size"])
    cam = R.camera(g["eye"].tolist(), g["target"].tolist(), g["up"].tolist(), float(g["vfov"]), w, h)
    tg = R.RayTargets(ctx, w, h, index_bytes=index_bytes, counters=counters)
    R.raymarch(ctx, sc, cam, flags | (L.RAY_COUNT if counters else 0), sh, tg)
    ctx.sync()
    return tg


def _assert_frame(tg, g, what):
    idx, face, t, shade = tg.results()
    assert np.array_equal(idx, g["idx"]), what
    assert np.array_equal(face, g["face"]), what
    assert t.tobytes() == g["t"].tobytes(), what
    assert shade.tobytes() == g["shade"].tobytes(), what


def test_spheres_fixture(ctx):
    g = _load("spheres")
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    assert np.array_equal(sc.arrays()["dense"], g["grid"].ravel())
    _assert_frame(_frame(ctx, sc, g, L.RAY_SHADOW, R.shading(LIGHT)), g, "spheres")
    sc.destroy()


def test_heightfield_fixture(ctx):
    g = _load("heightfield")
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, 64, 2)
    a = sc.arrays()
    assert np.array_equal(a["slots"], g["slots"])
    assert np.array_equal(a["masks"], g["masks"]) and np.array_equal(a["voxels"], g["voxels"])
    tg = _frame(ctx, sc, g, L.RAY_SHADOW, R.shading(LIGHT), counters=True)
    _assert_frame(tg, g, "heightfield")
    cnt = tg.counts()
    for k, v in zip(g["count_names"], g["counts"]):
        assert cnt[str(k)] == int(v), k
    sc.destroy()


def test_caves_fixture(ctx):
    g = _load("caves")
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, 256, 5)
    a = sc.arrays()
    assert len(a["masks"]) == int(g["n_bricks"])
    h = hashlib.sha256()
    for k in ("top", "slots", "masks", "voxels"):
        h.update(np.ascontiguousarray(a[k]).tobytes())
    assert h.digest() == g["scene_sha256"].tobytes()
    tg = _frame(ctx, sc, g, L.RAY_SHADOW | L.RAY_AO, R.shading(LIGHT, 9, 4, 16.0))
    _assert_frame(tg, g, "caves")
    sc.destroy()


def test_chunks_fixture(ctx):
    g = _load("chunks")
    n = len(g["voxels"])
    vox = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 3, n, vox.ptr, None))
    assert np.array_equal(vox.download(np.uint16).reshape(n, 32768), g["voxels"])
    src = R.chunk_source(vox, 2, n)
    for name, mode in (("culled", L.MESH_CULLED), ("greedy", L.MESH_GREEDY)):
        mt = R.MeshTargets(ctx, n, len(g[name + "_quads"]) + 16)
        R.mesh_chunks(ctx, src, mode, mt)
        ctx.sync()
        counts, prefix, total, quads = mt.results()
        assert np.array_equal(counts, g[name + "_counts"]), name
        assert total == len(g[name + "_quads"]) and np.array_equal(quads, g[name + "_quads"]), name


def test_edits_fixture(ctx):
    g = _load("edits")
    sc = ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, 200, 41)
    xyz = ctx.upload(g["xyz"].ravel())
    n = len(g["xyz"])
    R.edit_chunks(sc, xyz, n, 77)
    a = sc.arrays()
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(a["masks"]).tobytes())
    h.update(np.ascontiguousarray(a["voxels"]).tobytes())
    assert h.digest() == g["scene_sha256"].tobytes()
    mt = R.MeshTargets(ctx, n, len(g["quads"]) + 16)
    R.mesh_scene_chunks(ctx, sc, xyz, n, L.MESH_GREEDY, mt)
    ctx.sync()
    counts, prefix, total, quads = mt.results()
    assert np.array_equal(counts, g["counts"]) and np.array_equal(quads, g["quads"])
    sc.destroy()


def test_reference_sphere_fixture(ctx):
    g = _load("reference_sphere")
    nc = len(g["coords"])
    dc = ctx.upload(g["coords"].ravel())
    mips = ctx.buffer(nc * 4 * 4096)
    L.check(ctx.lib.meso_generate_sphere_chunks(ctx.h, dc.ptr, nc, 16, 1.0, 4, mips.ptr, None))
    ctx.sync()
    got = mips.download(np.uint8).reshape(nc, 4, 16, 16, 16)
    off = 0
    for i, k in enumerate(g["counts"]):
        cells = g["cells"][off:off + int(k)].astype(np.int64)
        off += int(k)
        occ = np.zeros((16, 16, 16), np.uint8)
        if len(cells):
            occ[cells[:, 2], cells[:, 1], cells[:, 0]] = 1
        assert np.array_equal(got[i, 0], occ), i
    assert np.array_equal(got[0], g["mips0"])
