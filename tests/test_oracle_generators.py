"""Pins of the oracle's seeded generators (CPU only)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O


def pcg_stream(seed, count):
    """PCG32 seeded through two SplitMix64 draws (pinned by the known answers)."""
    a, b = O.splitmix64_seq(seed, 2)
    r = O.pcg32_srandom(a, b)
    return [O.pcg32_next(r) for _ in range(count)]


def test_sphere_params_follow_the_draw_rule():
    n, k = 128, 8
    p = O.sphere_params(1, n, k)
    draws = pcg_stream(1, 4 * k)
    for i in range(k):
        assert p[i, 0] == draws[4 * i] % (n * 256)
        assert p[i, 3] == 8 * 256 + draws[4 * i + 3] % (24 * 256 + 1)


def test_spheres_dense_equals_numpy_lattice():
    n = 128
    p = O.sphere_params(1, n, 8).astype(np.int64)
    g = O.spheres_dense(1, n, 8)
    z, y, x = np.indices((n, n, n), dtype=np.int64)
    ref = np.zeros((n, n, n), np.uint8)
    for i in range(7, -1, -1):  # lowest index wins
        d2 = (x * 256 + 128 - p[i, 0]) ** 2 + (y * 256 + 128 - p[i, 1]) ** 2 + \
             (z * 256 + 128 - p[i, 2]) ** 2
        ref[d2 <= p[i, 3] ** 2] = i + 1
    assert np.array_equal(g, ref)
    assert 0.05 < (g > 0).mean() < 0.3


def test_value_noise_interpolates_lattice_hashes():
    L = O.lib()
    for (ix, iy, iz, s) in [(0, 0, 0, 1), (5, -3, 7, 99), (-8, 12, 3, 0xabc)]:
        h = L.mo_lattice_hash3(ix, iy, iz, s)
        assert L.mo_value_noise3(float(ix), float(iy), float(iz), s) == np.float32((h >> 8) / 2**24)
        # quintic fade: fade(1/2) = 1/2, so the cell centre on x is the corner mean
        h1 = L.mo_lattice_hash3(ix + 1, iy, iz, s)
        mid = L.mo_value_noise3(ix + 0.5, float(iy), float(iz), s)
        assert abs(mid - ((h >> 8) + (h1 >> 8)) / 2**25) < 1e-6
    # 2D noise at lattice points
    h = L.mo_hash32(3 ^ L.mo_hash32(4 ^ 17))
    assert L.mo_value_noise2(3.0, 4.0, 17) == np.float32((h >> 8) / 2**24)


def test_fbm_single_octave_is_noise_and_range():
    L = O.lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        x, y, z = (float(np.float32(v)) for v in rng.uniform(-50, 50, 3))
        assert L.mo_fbm3(x, y, z, 1, 0.25, 7) == L.mo_value_noise3(np.float32(x * 0.25),
                                                                  np.float32(y * 0.25),
                                                                  np.float32(z * 0.25), 7)
        v = L.mo_fbm3(x, y, z, 4, 0.125, 7)
        assert 0.0 <= v < 1.0


def test_heightfield_columns_and_fill():
    L = O.lib()
    n, seed = 512, 2
    for (x, z) in [(0, 0), (100, 200), (511, 511), (333, 17)]:
        h = L.mo_height(x, z, n, seed)
        col = [L.mo_scene_voxel(O.GEN_HEIGHTFIELD, x, y, z, n, seed) for y in range(n)]
        solid = [v for v in col if v]
        assert len(solid) == max(h, 0)
        for y, v in enumerate(col[:max(h, 0)]):
            assert v == 1 + min(3, (y * 16) // n)
    slots, masks, vox = O.build_brickmap(O.GEN_HEIGHTFIELD, n, seed)
    fill = np.unpackbits(masks.view(np.uint8)).sum() / n ** 3
    assert 0.08 < fill < 0.13


def test_brickmap_round_trip_and_mask_layout():
    n = 128
    slots, masks, vox = O.build_brickmap(O.GEN_TERRAIN_CAVES, n, 40)
    dense = O.dense_from_bricks(n, slots, vox)
    s2, m2, v2 = O.brickmap_from_dense(dense)
    assert np.array_equal(slots, s2) and np.array_equal(masks, m2) and np.array_equal(vox, v2)
    # mask word z, bit x + 8 y mirrors the voxel payload
    for b in range(0, len(masks), max(1, len(masks) // 50)):
        v = vox[b].reshape(8, 8, 8)
        for lz in range(8):
            bits = int(masks[b, lz])
            for ly in range(8):
                for lx in range(8):
                    assert ((bits >> (lx + 8 * ly)) & 1) == (v[lz, ly, lx] != 0)
    # slots are assigned in linear brick order
    used = slots[slots != O.EMPTY_SLOT]
    assert np.array_equal(used, np.arange(len(used)))


def test_gen_brick_matches_voxel_function():
    n, seed = 256, 5
    L = O.lib()
    rng = np.random.default_rng(1)
    for _ in range(6):
        bx, by, bz = (int(v) for v in rng.integers(0, n // 8, 3))
        s, vox, mask = O.gen_brick(O.GEN_CAVES, n, seed, bx, by, bz)
        for i in rng.integers(0, 512, 40):
            lx, ly, lz = i % 8, (i // 8) % 8, i // 64
            assert vox[i] == L.mo_scene_voxel(O.GEN_CAVES, bx * 8 + lx, by * 8 + ly, bz * 8 + lz,
                                              n, seed)
        assert s == int((vox > 0).sum())


def test_chunk_fill_is_beta_half_half():
    params = O.chunk_params(3, 4096)
    f = np.sort(np.array([p.fill for p in params], np.float64))
    # Kolmogorov-Smirnov against the arcsine CDF (2/pi) asin(sqrt(x))
    cdf = 2 / np.pi * np.arcsin(np.sqrt(f))
    n = len(f)
    ks = max(np.max(np.arange(1, n + 1) / n - cdf), np.max(cdf - np.arange(n) / n))
    assert ks < 1.63 / math.sqrt(n)  # p > 0.01
    counts = [p.n_spheres for p in params]
    assert set(counts) == {1, 2, 3, 4}


def test_chunk_noise_threshold_meets_target():
    params = O.chunk_params(8, 6)
    for p in params:
        p.n_spheres = 0  # noise only
        vox = np.zeros(32768, np.uint16)
        O.lib().mo_gen_chunk(C.cast(C.byref(p), C.c_void_p), vox.ctypes.data_as(C.c_void_p))
        target = int(np.float32(np.float32(p.fill * np.float32(32768.0)) + np.float32(0.5)))
        solid = int((vox > 0).sum())
        assert solid >= target
        assert set(np.unique(vox[vox > 0]).tolist()) <= {1, 2, 3, 4}


@pytest.mark.parametrize("n", [128, 200])
def test_scene_chunk_and_edits_match_dense_restatement(n):
    nc = -(-n // 32)  # a partial last chunk counts
    slots, masks, vox = O.build_brickmap(O.GEN_TERRAIN_CAVES, n, 41)
    masks, vox = masks.copy(), vox.copy()
    sc = O.brickmap_scene(n, slots, masks, vox)
    dense = O.dense_from_bricks(n, slots, vox)
    pad = np.zeros((nc * 32,) * 3, dense.dtype)
    pad[:n, :n, :n] = dense
    for c in [(0, 0, 0), (3, 1, 2), (1, 3, 3), (nc - 1, nc - 1, 0)]:
        assert np.array_equal(O.scene_chunk(sc, *c).reshape(32, 32, 32),
                              pad[c[2] * 32:(c[2] + 1) * 32, c[1] * 32:(c[1] + 1) * 32,
                                  c[0] * 32:(c[0] + 1) * 32])
    xyz = np.array([[0, 0, 0], [2, 1, 3], [3, 0, 1], [nc - 1, 1, nc - 1], [nc, 0, 0]], np.int32)
    O.edit_chunks(sc, xyz, 77)
    # restatement: toggle every voxel of an allocated brick inside the chunk's sphere
    H = O.lib().mo_hash32
    alloc = np.zeros((n, n, n), bool)
    s3 = slots.reshape(n // 8, n // 8, n // 8)
    for bz, by, bx in zip(*np.nonzero(s3 != O.EMPTY_SLOT)):
        alloc[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = True
    z, y, x = np.indices((n, n, n), dtype=np.int64)
    exp = dense.copy()
    for cx, cy, cz in xyz:
        if max(cx, cy, cz) >= nc:
            continue  # outside the chunk grid: skipped
        h0 = H(77 ^ H(int(cx + nc * (cy + nc * cz))))
        h1 = H(h0)
        h2 = H(h1)
        h3 = H(h2)
        sx, sy, sz = cx * 8192 + (h0 & 0x1FFF), cy * 8192 + (h1 & 0x1FFF), cz * 8192 + (h2 & 0x1FFF)
        r = 4 * 256 + h3 % (8 * 256)
        inside = ((x * 256 + 128 - sx) ** 2 + (y * 256 + 128 - sy) ** 2 + (z * 256 + 128 - sz) ** 2
                  <= r * r) & alloc & (x // 32 == cx) & (y // 32 == cy) & (z // 32 == cz)
        exp[inside] = np.where(exp[inside] != 0, 0, 7)
    assert np.array_equal(O.dense_from_bricks(n, slots, vox), exp)
    s2, m2, v2 = O.brickmap_from_dense(exp)
    # masks follow the voxels; bricks that lost every voxel keep their slot
    ref_m = np.zeros_like(masks)
    for b in range(len(vox)):
        v = vox[b].reshape(8, 64)
        ref_m[b] = [sum(1 << j for j in range(64) if v[lz, j]) for lz in range(8)]
    assert np.array_equal(masks, ref_m)


def test_two_level_node_major_order():
    n = 256
    slots, masks, vox = O.build_brickmap(O.GEN_CAVES, n, 5)
    top, ns, m2, v2 = O.two_level_from_bricks(n, slots, masks, vox)
    nodes = top[top != O.EMPTY_SLOT]
    assert np.array_equal(nodes, np.arange(len(nodes)))
    used = ns[ns != O.EMPTY_SLOT]
    assert np.array_equal(used, np.arange(len(used)))
    assert len(m2) == len(masks) == len(used)
