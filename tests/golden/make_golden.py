"""Writes the golden fixtures of the voxel path with the CPU oracle.

The reference has no test vectors for this path (it has no raymarcher or
mesher), so the fixtures are the oracle's outputs on small seeded inputs.
The oracle itself is pinned by tests/test_oracle_*.py (brute force, closed
forms, invariants, published PCG32 / SplitMix64 outputs); the fixtures
freeze those outputs so that any later drift of the oracle or of the GPU path
shows up as a fixture mismatch.

Each fixture is an .npz (numpy arrays only, loaded with allow_pickle=False);
SHA256SUMS lists their hashes.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

LIGHT = (0.4, 1.0, 0.3)


def _f32(v):
    return np.array(v, np.float32)


def spheres():
    """128^3 sphere grid (seed 1, 8 spheres) and a 64x48 frame with shadows."""
    grid = O.spheres_dense(1, 128, 8)
    cam = dict(eye=_f32([348.0, 230.0, 190.0]), target=_f32([64, 64, 64]), up=_f32([0, 0, 1]),
               vfov=np.float32(60.0), size=np.array([64, 48], np.int32))
    ocam = O.camera(cam["eye"].tolist(), cam["target"].tolist(), cam["up"].tolist(), 60.0, 64, 48)
    idx, face, t, shade, _ = O.raymarch(O.dense_scene(grid.ravel()), ocam, O.RAY_SHADOW,
                                        O.shading(LIGHT))
    return dict(grid=grid, idx=idx, face=face, t=t, shade=shade, **cam)


def heightfield():
    """64^3 heightfield brick map (seed 2) and a 96x64 frame with shadows."""
    n = 64
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, n, 2)
    cam = dict(eye=_f32([-8.0, 40.0, 32.0]), target=_f32([40.0, 4.0, 32.0]), up=_f32([0, 1, 0]),
               vfov=np.float32(60.0), size=np.array([96, 64], np.int32))
    ocam = O.camera(cam["eye"].tolist(), cam["target"].tolist(), cam["up"].tolist(), 60.0, 96, 64)
    idx, face, t, shade, cnt = O.raymarch(O.brickmap_scene(n, slots, masks, vox), ocam,
                                          O.RAY_SHADOW, O.shading(LIGHT))
    c = cnt.as_dict()
    return dict(slots=slots, masks=masks, voxels=vox, idx=idx, face=face, t=t, shade=shade,
                counts=np.array([c[k] for k in sorted(c)], np.int64),
                count_names=np.array(sorted(c)), **cam)


def caves():
    """256^3 two-level caves (seed 5): scene digest and a 64x48 frame with
    shadow and 4 AO rays (AO seed 9, radius 16)."""
    n = 256
    slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, 5)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    h = hashlib.sha256()
    for a in (top, ns, m2, v2):
        h.update(np.ascontiguousarray(a).tobytes())
    cam = dict(eye=_f32([128.3, 128.2, 128.1]), target=_f32([230.4, 115.2, 204.8]),
               up=_f32([0, 1, 0]), vfov=np.float32(60.0), size=np.array([64, 48], np.int32))
    ocam = O.camera(cam["eye"].tolist(), cam["target"].tolist(), cam["up"].tolist(), 60.0, 64, 48)
    flags = O.RAY_SHADOW | O.RAY_AO
    idx, face, t, shade, _ = O.raymarch(O.two_level_scene(n, top, ns, m2, v2), ocam, flags,
                                        O.shading(LIGHT, 9, 4, 16.0))
    return dict(scene_sha256=np.frombuffer(h.digest(), np.uint8), n_bricks=np.int64(len(m2)),
                n_nodes=np.int64(int((top != O.EMPTY_SLOT).sum())), idx=idx, face=face, t=t,
                shade=shade, **cam)


def chunks():
    """The first 6 chunks of the seeded 32^3 batch (seed 3) and their culled
    and greedy quads."""
    vox, _ = O.gen_chunks(3, 6)
    out = dict(voxels=vox)
    for name, greedy in (("culled", False), ("greedy", True)):
        qs = [O.mesh_chunk(vox[c], greedy) for c in range(len(vox))]
        out[name + "_counts"] = np.array([len(q) for q in qs], np.uint32)
        out[name + "_quads"] = np.concatenate(qs).astype(np.uint32)
    return out


def edits():
    """Seeded chunk edits (seed 77) on a 200^3 terrain-caves brick map (seed 41,
    a partial last chunk) and the greedy quads of the edited chunks."""
    n = 200
    slots, masks, vox = O.build_brickmap(O.GEN_TERRAIN_CAVES, n, 41)
    masks, vox = masks.copy(), vox.copy()
    sc = O.brickmap_scene(n, slots, masks, vox)
    xyz = np.array([[0, 0, 0], [2, 1, 3], [6, 1, 6], [3, 0, 1], [6, 6, 6]], np.int32)
    O.edit_chunks(sc, xyz, 77)
    h = hashlib.sha256()
    h.update(masks.tobytes())
    h.update(vox.tobytes())
    qs = [O.mesh_chunk(O.scene_chunk(sc, *map(int, c)), True) for c in xyz]
    return dict(xyz=xyz, scene_sha256=np.frombuffer(h.digest(), np.uint8),
                counts=np.array([len(q) for q in qs], np.uint32),
                quads=np.concatenate(qs).astype(np.uint32))


def reference_sphere():
    """The reference's GenerateSphere cells (GeneratorHelper.h:120-150) for 4
    chunks of 16^3 and the erosion mips of the first (Chunk.h:73-94)."""
    coords = np.array([[5, 0, 0], [6, -1, 0], [4, 1, -1], [7, 2, 2]], np.int32)
    cells = [O.ref_generate_sphere(*map(int, c)) for c in coords]
    occ = np.zeros((16, 16, 16), np.uint8)
    for x, y, z in cells[0]:
        occ[z, y, x] = 1
    return dict(coords=coords, counts=np.array([len(c) for c in cells], np.int32),
                cells=np.concatenate(cells).astype(np.uint8), mips0=O.ref_erode_mips(occ, 4))


FIXTURES = dict(spheres=spheres, heightfield=heightfield, caves=caves, chunks=chunks, edits=edits,
                reference_sphere=reference_sphere)


def main():
    lines = []
    for name, fn in FIXTURES.items():
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **fn())
        digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
        lines.append("%s  %s.npz\n" % (digest, name))
        print(name, os.path.getsize(path), digest)
    with open(os.path.join(HERE, "SHA256SUMS"), "w") as f:
        f.writelines(lines)


if __name__ == "__main__":
    main()
