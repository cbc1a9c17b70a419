"""Seeded chunk edits and meshing of chunks read straight from a resident
scene, against the oracle (the dirty-chunk remesh path of the 2048^3 frame)."""
import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_scene(kind, gen, n, seed):
    """oracle scene plus the (contiguous) mask and voxel arrays it edits in place"""
    slots, masks, vox = O.build_brickmap(gen, n, seed)
    if kind == L.SCENE_TWO_LEVEL_U16:
        top, slots, masks, vox = O.two_level_from_bricks(n, slots, masks, vox)
    masks, vox = np.ascontiguousarray(masks.copy()), np.ascontiguousarray(vox.copy())
    if kind == L.SCENE_BRICKMAP_U16:
        return O.brickmap_scene(n, slots, masks, vox), masks, vox
    return O.two_level_scene(n, top, slots, masks, vox), masks, vox


def _chunks(n, count, seed):
    nc = -(-n // 32)
    rng = np.random.default_rng(seed)
    flat = rng.choice(nc ** 3, size=min(count, nc ** 3), replace=False)
    return np.stack([flat % nc, (flat // nc) % nc, flat // (nc * nc)], 1).astype(np.int32)


@pytest.mark.parametrize("kind,gen,n,seed", [(L.SCENE_TWO_LEVEL_U16, L.GEN_CAVES, 256, 5),
                                             (L.SCENE_BRICKMAP_U16, L.GEN_TERRAIN_CAVES, 200, 8)])
def test_edit_then_remesh_from_scene(ctx, kind, gen, n, seed):
    sc = ctx.scene_generate(gen, kind, n, seed)
    osc, m_ref, v_ref = _oracle_scene(kind, gen, n, seed)
    xyz = _chunks(n, 60, 3)
    dxyz = ctx.upload(xyz.ravel())
    L.check(ctx.lib.meso_scene_edit_chunks(sc.h, dxyz.ptr, len(xyz), 1234, None))
    before = v_ref.copy()
    O.edit_chunks(osc, xyz, 1234)
    assert not np.array_equal(before, v_ref), "the edits changed nothing"
    a = sc.arrays()
    assert np.array_equal(a["masks"], m_ref) and np.array_equal(a["voxels"], v_ref)
    # remesh the edited chunks, plus one outside the scene
    xyz2 = np.concatenate([xyz, np.array([[99, 0, 0], [-1, 2, 2]], np.int32)])
    d2 = ctx.upload(xyz2.ravel())
    for mode in (L.MESH_GREEDY, L.MESH_CULLED):
        refs = [O.mesh_chunk(O.scene_chunk(osc, *map(int, c)), mode == L.MESH_GREEDY)
                if c.min() >= 0 and c.max() < -(-n // 32) else np.zeros(0, np.uint32) for c in xyz2]
        mt = R.MeshTargets(ctx, len(xyz2), sum(len(r) for r in refs) + 1)
        L.check(ctx.lib.meso_mesh_scene_chunks(ctx.h, sc.h, d2.ptr, len(xyz2), mode, mt.counts.ptr,
                                               mt.prefix.ptr, mt.quads.ptr, mt.capacity,
                                               mt.total.ptr, mt.work.ptr, None))
        ctx.sync()
        counts, prefix, total, quads = mt.results()
        for k, r in enumerate(refs):
            assert counts[k] == len(r), (mode, k, counts[k], len(r))
            assert np.array_equal(quads[prefix[k]:prefix[k] + counts[k]], r), (mode, k)
    # the edited scene still renders bit-exact, with shadow and AO rays
    cp = dict(eye=[-0.2 * n, 0.9 * n, -0.1 * n], target=[0.5 * n, 0.3 * n, 0.5 * n],
              up=[0.0, 1.0, 0.0], vfov=60.0)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], 60.0, 128, 96)
    ocam = O.camera(cp["eye"], cp["target"], cp["up"], 60.0, 128, 96)
    tg = R.RayTargets(ctx, 128, 96)
    flags = L.RAY_SHADOW | L.RAY_AO
    R.raymarch(ctx, sc, cam, flags, R.shading(W.LIGHT, 3, 4, 16.0), tg)
    ctx.sync()
    ref = O.raymarch(osc, ocam, flags, O.shading(W.LIGHT, 3, 4, 16.0), threads=8)
    for x, y in zip(tg.results(), ref[:4]):
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8))
    sc.destroy()
