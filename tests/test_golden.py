"""The committed golden fixtures (tests/golden/*.npz) are intact and the
oracle still reproduces every array in them."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_golden as G  # noqa: E402


def _manifest():
    out = {}
    for line in open(os.path.join(HERE, "SHA256SUMS")):
        digest, name = line.split()
        out[name] = digest
    return out


def test_manifest_matches_files():
    m = _manifest()
    assert sorted(m) == sorted(n + ".npz" for n in G.FIXTURES)
    for name, digest in m.items():
        assert hashlib.sha256(open(os.path.join(HERE, name), "rb").read()).hexdigest() == digest


@pytest.mark.parametrize("name", sorted(G.FIXTURES))
def test_oracle_reproduces_fixture(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        stored = {k: z[k] for k in z.files}
    fresh = G.FIXTURES[name]()
    assert sorted(stored) == sorted(fresh)
    for k, v in fresh.items():
        v = np.asarray(v)
        assert stored[k].dtype == v.dtype and stored[k].shape == v.shape, (name, k)
        if v.dtype.kind == "f":  # floats bit for bit
            assert stored[k].tobytes() == v.tobytes(), (name, k)
        else:
            assert np.array_equal(stored[k], v), (name, k)
