"""Pins of the oracle's culled-face and greedy meshers and of the reference's
erosion cull (CPU only)."""
import numpy as np
import pytest

import oracle as O

FACE_AXIS = {0: 2, 1: 2, 2: 1, 3: 1, 4: 0, 5: 0}
FACE_SIGN = {0: 1, 1: -1, 2: 1, 3: -1, 4: 1, 5: -1}


def exposed_faces(v):
    """Brute force: set of (x, y, z, face) with a solid cell and an empty (or
    outside) neighbour across the face.  v: [z][y][x]."""
    n = v.shape[0]
    pad = np.zeros((n + 2,) * 3, v.dtype)
    pad[1:-1, 1:-1, 1:-1] = v
    out = set()
    for f in range(6):
        a, s = FACE_AXIS[f], FACE_SIGN[f]
        off = [0, 0, 0]
        off[2 - a] = s  # array axes are [z][y][x]
        nb = pad[1 + off[0]:n + 1 + off[0], 1 + off[1]:n + 1 + off[1], 1 + off[2]:n + 1 + off[2]]
        zs, ys, xs = np.nonzero((v != 0) & (nb == 0))
        out |= {(int(x), int(y), int(z), f) for x, y, z in zip(xs, ys, zs)}
    return out


def quad_cells(q):
    """cells (x, y, z, face) covered by packed quads"""
    u = O.unpack_quads(q)
    cells = []
    for i in range(len(q)):
        x, y, z, w, h, f = (int(u[k][i]) for k in ("x", "y", "z", "w", "h", "face"))
        for j in range(h):
            for k in range(w):
                if f < 2:
                    cells.append((x + k, y + j, z, f))
                elif f < 4:
                    cells.append((x + k, y, z + j, f))
                else:
                    cells.append((x, y + k, z + j, f))
    return cells


def chunk(v):
    return np.ascontiguousarray(v.reshape(-1))


def test_culled_equals_brute_force_and_order():
    vox, _ = O.gen_chunks(21, 6)
    for c in range(6):
        v = vox[c].reshape(32, 32, 32)
        q = O.mesh_chunk(vox[c], False)
        u = O.unpack_quads(q)
        assert (u["w"] == 1).all() and (u["h"] == 1).all() and (u["pad"] == 0).all()
        cells = quad_cells(q)
        assert len(cells) == len(set(cells))
        assert set(cells) == exposed_faces(v)
        # canonical order: face, slice, v, u
        f = u["face"].astype(np.int64)
        sl = np.where(f < 2, u["z"], np.where(f < 4, u["y"], u["x"])).astype(np.int64)
        vv = np.where(f < 2, u["y"], u["z"]).astype(np.int64)
        uu = np.where(f < 4, u["x"], u["y"]).astype(np.int64)
        key = ((f * 32 + sl) * 32 + vv) * 32 + uu
        assert (np.diff(key) > 0).all()


def test_greedy_tiles_the_culled_faces():
    vox, _ = O.gen_chunks(22, 8)
    for c in range(8):
        v = vox[c].reshape(32, 32, 32)
        q = O.mesh_chunk(vox[c], True)
        cells = quad_cells(q)
        assert len(cells) == len(set(cells)), "quads overlap"
        assert set(cells) == exposed_faces(v), "quads do not tile the exposed faces"
        # one material and one face id per quad
        u = O.unpack_quads(q)
        start = 0
        for i in range(len(q)):
            h, w = int(u["h"][i]), int(u["w"][i])
            cs = cells[start:start + w * h]
            start += w * h
            mats = {int(v[z, y, x]) for x, y, z, _ in cs}
            assert len(mats) == 1
        assert len(q) <= len(O.mesh_chunk(vox[c], False))


def test_single_voxel_and_box_closed_forms():
    v = np.zeros((32, 32, 32), np.uint16)
    v[5, 6, 7] = 3
    assert len(O.mesh_chunk(chunk(v), False)) == 6
    assert len(O.mesh_chunk(chunk(v), True)) == 6
    for (a, b, c) in [(1, 1, 1), (3, 4, 5), (32, 32, 32), (32, 1, 7)]:
        v = np.zeros((32, 32, 32), np.uint16)
        v[:c, :b, :a] = 2  # x extent a, y extent b, z extent c
        culled = O.mesh_chunk(chunk(v), False)
        greedy = O.mesh_chunk(chunk(v), True)
        assert len(culled) == 2 * (a * b + b * c + c * a)
        assert len(greedy) == 6
        u = O.unpack_quads(greedy)
        dims = sorted(zip(u["face"].tolist(), u["w"].tolist(), u["h"].tolist()))
        assert dims == sorted([(0, a, b), (1, a, b), (2, a, c), (3, a, c), (4, b, c), (5, b, c)])


def test_materials_split_quads():
    v = np.zeros((32, 32, 32), np.uint16)
    v[:4, :4, :8] = 1
    v[:4, :4, 4:8] = 2  # the same box in two materials along x
    q = O.mesh_chunk(chunk(v), True)
    assert len(q) == 10  # each half: 5 exposed sides, the shared wall is interior


def test_checkerboard_worst_case():
    z, y, x = np.indices((32, 32, 32))
    v = (((x + y + z) % 2) == 0).astype(np.uint8)
    q = O.mesh_chunk(chunk(v), True)
    assert len(q) == 16384 * 6 == 98304


def test_u8_and_u16_agree():
    vox, _ = O.gen_chunks(23, 2)
    v8 = np.minimum(vox[0], 255).astype(np.uint8)
    assert np.array_equal(O.mesh_chunk(v8, True), O.mesh_chunk(v8.astype(np.uint16), True))


def test_erosion_cull_has_no_exposed_faces():
    # cells the reference culls (mip 1 of the erosion chain, ChunkPool.h:385-390)
    # have no exposed face in the culled-face mesh of the same 16^3 chunk
    rng = np.random.default_rng(3)
    for trial in range(4):
        occ = np.zeros((16, 16, 16), np.uint8)
        for _ in range(6):
            c = rng.integers(0, 16, 3)
            r = rng.integers(2, 7)
            occ[max(0, c[2] - r):c[2] + r, max(0, c[1] - r):c[1] + r, max(0, c[0] - r):c[0] + r] = 1
        mips = O.ref_erode_mips(occ, 4)
        v = np.zeros((32, 32, 32), np.uint8)
        v[:16, :16, :16] = occ
        faces = exposed_faces(v)
        exposed_cells = {(x, y, z) for x, y, z, _ in faces}
        zs, ys, xs = np.nonzero(mips[1])
        assert len(xs) > 0
        assert not ({(int(x), int(y), int(z)) for x, y, z in zip(xs, ys, zs)} & exposed_cells)
        # mips shrink: mip d is a subset of mip d-1
        for d in range(1, 4):
            assert not (mips[d] & ~mips[d - 1]).any()


def test_erosion_of_full_block_closed_form():
    occ = np.ones((16, 16, 16), np.uint8)
    mips = O.ref_erode_mips(occ, 4)
    # cells within 1 of the border are never set above mip 0
    # (bIsOutOfBoundThickness), so each level peels one more layer
    assert [int(m.sum()) for m in mips] == [4096, 14 ** 3, 12 ** 3, 10 ** 3]
    occ = np.zeros((16, 16, 16), np.uint8)
    occ[2:14, 2:14, 2:14] = 1
    mips = O.ref_erode_mips(occ, 4)
    assert [int(m.sum()) for m in mips] == [12 ** 3, 10 ** 3, 8 ** 3, 6 ** 3]


def test_imposter_faces_cover_visible_culled_faces():
    # The reference draws each cell as the 3 faces selected by the octant of
    # (cell - camera) (GetOctantId, SimpleVoxel.cpp:129-136).  For a camera
    # outside the chunk, every culled face the camera can see is one of them.
    rng = np.random.default_rng(4)
    vox, _ = O.gen_chunks(24, 1)
    v = vox[0].reshape(32, 32, 32)
    faces = exposed_faces(v)
    L = O.lib()
    for _ in range(5):
        cam = rng.normal(size=3)
        cam = 16 + cam / np.linalg.norm(cam) * 60
        for (x, y, z, f) in list(faces)[:3000]:
            centre = np.array([x, y, z]) + 0.5
            a, s = FACE_AXIS[f], FACE_SIGN[f]
            facing = (cam[a] - (centre[a] + 0.5 * s)) * s > 0
            if not facing:
                continue
            rel = centre - cam
            oc = L.mo_ref_octant_id(float(rel[0]), float(rel[1]), float(rel[2]))
            # octant bit set => component negative => the camera is on the + side
            plus = [(oc >> k) & 1 == 1 for k in range(3)]
            assert plus[a] == (s > 0)


def test_reference_sphere_generator_lattice_count():
    # GenerateSphere with block size 1: solid iff (x-100)^2 + y^2 + z^2 < 2500
    # at integer corners; the count over chunks covering the ball equals the
    # lattice count, and cells come X-outermost, Z-innermost.
    total = 0
    for cx in range(3, 10):
        for cy in range(-4, 4):
            for cz in range(-4, 4):
                cells = O.ref_generate_sphere(cx, cy, cz, 1.0, 16)
                total += len(cells)
                if len(cells):
                    keys = cells[:, 0].astype(np.int64) * 256 + cells[:, 1] * 16 + cells[:, 2]
                    assert (np.diff(keys) > 0).all()
    xs = np.arange(48, 160)[:, None, None]
    ys = np.arange(-64, 64)[None, :, None]
    zs = np.arange(-64, 64)[None, None, :]
    lattice = int(((xs - 100) ** 2 + ys ** 2 + zs ** 2 < 2500).sum())
    assert total == lattice


@pytest.mark.parametrize("greedy", [False, True])
def test_mesh_chunks_mt_counts_equal_serial(greedy):
    vox, _ = O.gen_chunks(25, 16)
    counts, total = O.mesh_chunks_counts(vox, greedy, threads=4)
    serial = [len(O.mesh_chunk(vox[c], greedy)) for c in range(16)]
    assert counts.tolist() == serial and total == sum(serial)
