"""The oracle's restatement of the reference's noise, pinned by what the
mathematics fixes.

FVoxelMathHelper::Hash (Runtimes/Helper/VoxelMathHelper.h:29-33),
FGeneratorHelper::noised / displacement / TestGenerator
(Runtimes/Helper/GeneratorHelper.h:18-119).  These feed only the CPU-baseline
timing of the reference's own chunk generator; their exact bits follow the
platform libm sin.  What is fixed independently of libm:

- Hash(0) = fract(sin(0) * k) = 0, values lie in [0, 1), and the dot product
  weights x by 127.1, y by 311.7 and z by 74.7 (a swapped pair of weights
  changes which coordinate permutations collide).
- noised is a quintic-faded trilinear interpolation of 2 Hash - 1 at the
  lattice: at a lattice point it equals 2 Hash(p) - 1 with a zero gradient
  (the fade's derivative vanishes there), at a cell centre the fade is 1/2 so
  the value is 2 * mean(8 corner hashes) - 1, the function is continuous
  across cell faces, and its analytic gradient equals finite differences.
- displacement sums 5 domain-warped octaves and 4 plain octaves with
  amplitudes 1/2 .. 1/512, so |displacement| < 1; composed here from the
  pinned noised it must give the oracle's value.
- TestGenerator's test (y/2 + 10.3 d) * 0.4 < 0 with |d| < 1 makes every
  cell with y < -20.6 solid and every cell with y > 20.6 empty.
"""
import numpy as np
import pytest

import oracle as O


def test_hash_zero_range_and_axis_weights():
    assert O.ref_hash3(0.0, 0.0, 0.0) == 0.0
    rng = np.random.default_rng(1)
    for p in rng.integers(-5000, 5000, size=(500, 3)):
        h = O.ref_hash3(*p)
        assert 0.0 <= h < 1.0
    # x * 127.1 with x = 311.7 is the same product as y * 311.7 with y = 127.1
    assert O.ref_hash3(311.7, 0, 0) == O.ref_hash3(0, 127.1, 0)
    assert O.ref_hash3(74.7, 0, 0) == O.ref_hash3(0, 0, 127.1)
    assert O.ref_hash3(0, 74.7, 0) == O.ref_hash3(0, 0, 311.7)
    # and not for the other pairings
    assert O.ref_hash3(127.1, 0, 0) != O.ref_hash3(0, 311.7, 0)


def _corners(p):
    return [O.ref_hash3(p[0] + i, p[1] + j, p[2] + k)
            for k in (0, 1) for j in (0, 1) for i in (0, 1)]


@pytest.mark.parametrize("p", [(0, 0, 0), (3, -7, 12), (-40, 5, -1), (1000, 2000, -3000)])
def test_noised_at_lattice_points(p):
    r = O.ref_noised(np.array(p, np.float64))
    assert r[3] == -1.0 + 2.0 * O.ref_hash3(*p)
    assert r[0] == 0.0 and r[1] == 0.0 and r[2] == 0.0


@pytest.mark.parametrize("p", [(0, 0, 0), (3, -7, 12), (-40, 5, -1), (77, 78, 79)])
def test_noised_cell_centre_is_mean_of_corners(p):
    r = O.ref_noised(np.array(p, np.float64) + 0.5)
    assert abs(r[3] - (2.0 * np.mean(_corners(p)) - 1.0)) < 1e-12


def test_noised_continuous_across_cell_faces():
    rng = np.random.default_rng(2)
    for _ in range(200):
        x = rng.uniform(-50, 50, 3)
        axis = int(rng.integers(3))
        x[axis] = np.floor(x[axis]) + 1.0
        lo, hi = x.copy(), x.copy()
        lo[axis] -= 1e-9
        assert abs(O.ref_noised(lo)[3] - O.ref_noised(hi)[3]) < 1e-7


def test_noised_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(300):
        x = rng.uniform(-100, 100, 3)
        g = O.ref_noised(x)[:3]
        for a in range(3):
            e = np.zeros(3)
            e[a] = h
            fd = (O.ref_noised(x + e)[3] - O.ref_noised(x - e)[3]) / (2 * h)
            assert abs(fd - g[a]) < 1e-6 * max(1.0, abs(g[a])), (x, a, fd, g[a])


def _displacement(p0):
    """the octave composition of GeneratorHelper.h:55-87 over the pinned noised"""
    p = np.array(p0, np.float64)
    d, mgn, s = 0.0, 0.5, 1.0
    for i in range(5):
        r = O.ref_noised(p + 10.0)
        d += r[3] * mgn
        p = p * 2.0 + r[:3] * 0.2 * s
        if i == 2:
            s = -s
        mgn *= 0.5
    p = np.array(p0, np.float64) * 32.0
    for _ in range(4):
        r = O.ref_noised(p)
        d += r[3] * mgn
        p = p * 2.0
        mgn *= 0.5
    return d


def test_displacement_octaves_and_bound():
    rng = np.random.default_rng(4)
    for p in rng.uniform(-30, 30, size=(200, 3)):
        d = O.ref_displacement(p)
        assert abs(d - _displacement(p)) < 1e-12
        assert abs(d) < 1.0 - 2.0 ** -9 + 1e-12


def test_test_generator_height_bounds():
    # chunk y in [-48, -33]: y / 2 <= -16.5 < -10.3 |d|, every cell solid
    assert len(O.ref_test_generator(0, -3, 0)) == 16 ** 3
    assert len(O.ref_test_generator(7, -5, -2)) == 16 ** 3
    # chunk y in [32, 47]: every cell empty
    assert len(O.ref_test_generator(0, 2, 0)) == 0
    assert len(O.ref_test_generator(-4, 3, 9)) == 0
    # a chunk through the surface: the listed cells are the ones the test
    # keeps, in X-outer, Z-inner order, at cell corners chunk * 16 + (X, Y, Z)
    cx, cy, cz = 2, -1, 3
    cells = O.ref_test_generator(cx, cy, cz)
    assert 0 < len(cells) < 16 ** 3
    keys = [tuple(int(v) for v in c) for c in cells]
    assert keys == sorted(keys)
    want = []
    for X in range(16):
        for Y in range(16):
            for Z in range(16):
                p = np.array([cx * 16 + X, cy * 16 + Y, cz * 16 + Z], np.float64)
                if (p[1] * 0.5 + O.ref_displacement(p * 0.1) * 10.3) * 0.4 < 0.0:
                    want.append((X, Y, Z))
    assert keys == want
