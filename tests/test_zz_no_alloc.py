"""meso_mesh_chunks_noalloc against meso_mesh_chunks, through the C-ABI.

Two calls in a row on the same GPU (the second with fewer chunks, so a
scratch the first call left behind would show) must give the quads, counts,
prefix and total of the entry point that takes a workspace, for both mesh
modes; a batch above the scratch's maximum is refused.
"""
import numpy as np
import pytest

from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R

pytestmark = pytest.mark.gpu

MAX = 16384


def _mesh(ctx, src, mode, n, total_cap, noalloc):
    mt = R.MeshTargets(ctx, n, total_cap)
    # stale contents: a call must write every output it reports
    for b in (mt.counts, mt.prefix, mt.total):
        b.fill(0xEE)
    if noalloc:
        L.check(ctx.lib.meso_mesh_chunks_noalloc(ctx.h, L.byref(src), mode, mt.counts.ptr,
                                                 mt.prefix.ptr, mt.quads.ptr, mt.capacity,
                                                 mt.total.ptr, None))
    else:
        R.mesh_chunks(ctx, src, mode, mt)
    ctx.sync()
    return mt.results()


@pytest.mark.parametrize("mode", [L.MESH_GREEDY, L.MESH_CULLED])
def test_noalloc_equals_workspace_entry_point(ctx, mode):
    n0 = 600
    vox = ctx.buffer(2 * 32768 * n0)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 21, n0, vox.ptr, None))
    for n in (n0, 150):  # the second call is smaller
        src = R.chunk_source(vox, 2, n)
        ref = _mesh(ctx, src, mode, n, 6_000_000, False)
        got = _mesh(ctx, src, mode, n, 6_000_000, True)
        assert ref[2] > 0 and got[2] == ref[2], (n, got[2], ref[2])
        assert np.array_equal(got[0], ref[0]), n
        assert np.array_equal(got[1], ref[1]), n
        assert np.array_equal(got[3], ref[3]), n


def test_noalloc_refuses_batches_above_its_scratch(ctx):
    vox = ctx.buffer(2 * 32768)
    src = R.chunk_source(vox, 2, MAX + 1)
    mt = R.MeshTargets(ctx, 1, 16)
    st = ctx.lib.meso_mesh_chunks_noalloc(ctx.h, L.byref(src), L.MESH_GREEDY, mt.counts.ptr,
                                          mt.prefix.ptr, mt.quads.ptr, mt.capacity, mt.total.ptr,
                                          None)
    assert st == L.ERR_ARG
    assert b"16384" in ctx.lib.meso_last_error_message()
