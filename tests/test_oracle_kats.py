"""Known-answer and convention pins of the oracle (CPU only)."""
import numpy as np

import oracle as O


def test_pcg32_published_outputs():
    # pcg32-demo (pcg-random.org, pcg-c-basic): pcg32_srandom_r(42, 54)
    r = O.pcg32_srandom(42, 54)
    got = [O.pcg32_next(r) for _ in range(6)]
    assert got == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


def test_splitmix64_published_outputs():
    # SplitMix64 (Vigna) from state 0
    assert O.splitmix64_seq(0, 3) == [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f]


def _lowbias32_inverse(x):
    # published inverse of lowbias32 (hash-prospector)
    M = 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x43021123) & M
    x ^= (x >> 15) ^ (x >> 30)
    x = (x * 0x1d69e2a5) & M
    x ^= x >> 16
    return x


def test_hash32_is_lowbias32():
    rng = np.random.default_rng(0)
    for x in rng.integers(0, 2**32, 2000, dtype=np.uint64):
        x = int(x)
        assert _lowbias32_inverse(O.lib().mo_hash32(x)) == x
    assert O.lib().mo_hash32(0) == 0


def test_lattice_hash_composition():
    L = O.lib()
    for (x, y, z, s) in [(0, 0, 0, 0), (1, -2, 3, 77), (-5, 9, 1 << 20, 0xdeadbeef)]:
        h = L.mo_hash32((x & 0xFFFFFFFF) ^ L.mo_hash32((y & 0xFFFFFFFF) ^ L.mo_hash32((z & 0xFFFFFFFF) ^ s)))
        assert L.mo_lattice_hash3(x, y, z, s) == h


def test_quad_pack_layout():
    q = O.lib().mo_quad_pack(31, 1, 2, 32, 3, 5)
    u = O.unpack_quads(np.array([q], np.uint32))
    assert (u["x"][0], u["y"][0], u["z"][0], u["w"][0], u["h"][0], u["face"][0], u["pad"][0]) == \
        (31, 1, 2, 32, 3, 5, 0)
    assert q == 31 | (1 << 5) | (2 << 10) | (31 << 15) | (2 << 20) | (5 << 25)


def _one_voxel(n, x, y, z):
    g = np.zeros((n, n, n), np.uint8)
    g[z, y, x] = 9
    return O.dense_scene(g)


def test_face_ids_follow_cube_vertex_groups():
    # Cube.h:25-54 lists the faces +Z, -Z, +Y, -Y, +X, -X; a ray entering the
    # voxel through its +Z face (travelling -z) must report face 0, and so on.
    n = 8
    sc = _one_voxel(n, 3, 4, 5)
    c = np.array([3.5, 4.5, 5.5], np.float32)
    cases = [((0, 0, -1), 0), ((0, 0, 1), 1), ((0, -1, 0), 2), ((0, 1, 0), 3), ((-1, 0, 0), 4),
             ((1, 0, 0), 5)]
    for d, face in cases:
        d = np.array(d, np.float32)
        o = (c - 20.0 * d).astype(np.float32)
        h, _ = O.trace(sc, o, d)
        assert h.face == face, (d, h.face)
        assert h.index == 3 + n * 4 + n * n * 5  # Convert3DTo1D, x fastest
        assert h.material == 9


def test_index_linearisation_x_fastest():
    n = 16
    for (x, y, z) in [(0, 0, 0), (15, 0, 0), (0, 15, 0), (0, 0, 15), (7, 3, 11)]:
        sc = _one_voxel(n, x, y, z)
        h, _ = O.trace(sc, np.array([x + 0.5, y + 0.5, -3.0], np.float32),
                       np.array([0, 0, 1], np.float32))
        assert h.index == x + y * n + z * n * n


def test_octant_id_matches_reference_rule():
    # GetOctantId (Samples/SimpleVoxel.cpp:129-136): bit 0 x<0, bit 1 y<0, bit 2 z<0
    L = O.lib()
    for x in (-1.0, 1.0):
        for y in (-1.0, 1.0):
            for z in (-1.0, 1.0):
                assert L.mo_ref_octant_id(x, y, z) == (x < 0) | ((y < 0) << 1) | ((z < 0) << 2)
    assert L.mo_ref_octant_id(0.0, 0.0, 0.0) == 0
