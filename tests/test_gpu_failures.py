"""Failure handling on the GPU, every condition forced from the host so that
every kernel still exits normally.

- The batched mesher's look-back wait is bounded; with its spin limit set to
  0 by the host, every chunk after the first gives up at its first look, the
  persistent launch still completes, and meso_mesh_check reports the error.
  The reference re-queues a chunk task whose dispatch fails
  (Runtimes/Voxel/Chunk/ChunkManager.h:355-372); here the failure is reported
  and the next batch, at the default limit, is exact again.
- Allocations that cannot succeed (buffers, a server's arena) return
  OUT_OF_MEMORY and leave the context usable.
- Stopping a server with requests still waiting for budget ends their waits
  with RUNTIME_ERROR instead of hanging.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_lookback_wait_gives_up_and_is_reported(ctx):
    n = 256
    vox = ctx.buffer(2 * 32768 * n)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, 31, n, vox.ptr, None))
    src = R.chunk_source(vox, 2, n)
    mt = R.MeshTargets(ctx, n, 4_000_000)
    try:
        L.check(ctx.lib.meso_mesh_set_lookback_spin_limit(ctx.h, 0))
        R.mesh_chunks(ctx, src, L.MESH_GREEDY, mt)
        ctx.sync()  # the launch completed: every wave left
        st = ctx.lib.meso_mesh_check(ctx.h, mt.work.ptr, None)
        assert st == L.ERR_RUNTIME
        assert b"gave up" in ctx.lib.meso_last_error_message()
    finally:
        L.check(ctx.lib.meso_mesh_set_lookback_spin_limit(ctx.h, 1 << 24))
    # the default bound: exact again, and the check passes
    R.mesh_chunks(ctx, src, L.MESH_GREEDY, mt)
    L.check(ctx.lib.meso_mesh_check(ctx.h, mt.work.ptr, None))
    counts, prefix, total, quads = mt.results()
    ref_vox, _ = O.gen_chunks(31, n)
    ref_counts, ref_total = O.mesh_chunks_counts(ref_vox, True, threads=8)
    assert np.array_equal(counts, ref_counts) and total == ref_total
    for c in range(0, n, 37):
        assert np.array_equal(quads[prefix[c]:prefix[c] + counts[c]],
                              O.mesh_chunk(ref_vox[c], True)), c


def test_impossible_allocations_leave_the_context_usable(ctx):
    h = C.c_void_p()
    assert ctx.lib.meso_buffer_create(ctx.h, 1 << 50, None, C.byref(h)) == L.ERR_OOM
    cfg = L.ServeConfig(C.sizeof(L.ServeConfig), 1 << 50, 8, 8, 0, 8, 0)
    assert ctx.lib.meso_serve_start(ctx.h, L.byref(cfg), C.byref(h)) == L.ERR_OOM
    # still usable
    sc = ctx.scene_generate(L.GEN_SPHERES, L.SCENE_DENSE_U8, 128, 1, 8)
    cp = W.spheres_camera(width=64, height=48)
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 64, 48)
    tg = R.RayTargets(ctx, 64, 48)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    ctx.sync()
    ref = O.raymarch(O.dense_scene(O.spheres_dense(1, 128, 8).ravel()),
                     O.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], 64, 48),
                     O.RAY_SHADOW, O.shading(W.LIGHT))
    assert np.array_equal(tg.results()[0], ref[0])
    sc.destroy()


def test_stop_with_requests_waiting_for_budget(ctx):
    sc = ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, 128, 40)
    sh = R.shading(W.LIGHT)
    # 640x360 needs ~3.2 MB: the 4 MiB budget holds one result at a time
    srv = R.Server(ctx, 4 << 20, max_pending=8, max_inflight=4)
    cam = R.camera((300.0, 200.0, 250.0), (64.0, 20.0, 64.0), (0, 1, 0), 60.0, 640, 360)
    held = srv.submit(sc, cam, L.RAY_SHADOW, sh)
    srv.wait(held, 30_000_000)  # complete, not taken: its bytes stay held
    waiting = srv.submit(sc, cam, L.RAY_SHADOW, sh)
    with pytest.raises(L.MesoError) as e:
        srv.wait(waiting, 50_000)
    assert e.value.status == L.ERR_BUSY  # times out: the budget binds
    srv.stop()  # returns with a request still waiting for budget
    sc.destroy()
