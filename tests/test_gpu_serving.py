"""Serving: continuous batching keeps every result identical alone and in a mix,
admission is bounded by the HBM budget and by the entries of the device
ring, back-pressure is reported, and the persistent kernel picks up work
published while it runs and is launched again after it idles out."""
import time

import numpy as np
import pytest

import oracle as O
from mesoengine_amd import _lib as L
from mesoengine_amd import runtime as R
from mesoengine_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(ctx):
    ss = [ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, e, 40 + i)
          for i, e in enumerate((128, 200, 256))]
    yield ss
    for s in ss:
        s.destroy()


def _req(edge, w, h, k):
    rng = np.random.default_rng(k)
    v = rng.normal(size=3)
    v[1] = abs(v[1]) + 0.3
    v /= np.linalg.norm(v)
    c = np.array([edge / 2, edge * 0.15, edge / 2])
    eye = c + 1.5 * edge * np.sqrt(3) * v
    return R.camera(np.float32(eye), np.float32(c), (0, 1, 0), 60.0, w, h)


def test_alone_equals_mix_and_plain(ctx, scenes):
    sh = R.shading(W.LIGHT)
    target = (1, 640, 360, 7)
    cam = _req(200, 640, 360, 7)
    srv = R.Server(ctx, 256 << 20, max_pending=64, max_inflight=8, idle_exit_us=200000)
    t = srv.submit(scenes[1], cam, L.RAY_SHADOW, sh)
    srv.wait(t, 30_000_000)
    alone = srv.take(t, 640, 360)
    # the same request in the middle of a mix of other sizes and scenes
    mix = []
    sizes = [(256, 256), (1280, 720), (640, 360), (1920, 1080), (256, 256)]
    for k in range(10):
        s = k % 3
        w, h = sizes[k % len(sizes)]
        mix.append((srv.submit(scenes[s], _req((128, 200, 256)[s], w, h, 100 + k), L.RAY_SHADOW, sh), w, h))
        if k == 4:
            tt = srv.submit(scenes[1], cam, L.RAY_SHADOW, sh)
    srv.wait(tt, 30_000_000)
    inmix = srv.take(tt, 640, 360)
    for tk, w, h in mix:
        srv.wait(tk, 30_000_000)
        srv.take(tk, w, h)
    st = srv.stats()
    srv.stop()
    for a, b in zip(alone, inmix):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    # and identical to the plain per-request path
    tg = R.RayTargets(ctx, 640, 360)
    R.raymarch(ctx, scenes[1], cam, L.RAY_SHADOW, sh, tg)
    ctx.sync()
    for a, b in zip(alone, tg.results()):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert st["completed"] == 12 and st["launches"] >= 1


def test_served_matches_oracle(ctx, scenes):
    sh = R.shading(W.LIGHT)
    cam = _req(128, 160, 120, 3)
    srv = R.Server(ctx, 64 << 20)
    t = srv.submit(scenes[0], cam, L.RAY_SHADOW, sh)
    srv.wait(t, 30_000_000)
    got = srv.take(t, 160, 120)
    srv.stop()
    a = scenes[0].arrays()
    osc = O.brickmap_scene(128, a["slots"], a["masks"], a["voxels"])
    # the scene itself equals the oracle's generator
    slots, masks, vox = O.build_brickmap(O.GEN_TERRAIN_CAVES, 128, 40)
    assert np.array_equal(a["slots"], slots) and np.array_equal(a["voxels"], vox)
    ocam = O.camera(list(cam.eye), list(cam.target), list(cam.up), 60.0, 160, 120)
    ref = O.raymarch(osc, ocam, O.RAY_SHADOW, O.shading(W.LIGHT))
    assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
    assert np.array_equal(got[2].view(np.uint32), ref[2].view(np.uint32))
    assert np.array_equal(got[3].view(np.uint32), ref[3].view(np.uint32))


def test_submit_rejects_bad_cameras(ctx, scenes):
    sh = R.shading(W.LIGHT)
    srv = R.Server(ctx, 16 << 20)
    try:
        bad = [R.camera((float("nan"), 0, 0), (1, 1, 1), (0, 1, 0), 60.0, 64, 64),
               R.camera((0, 0, 0), (1, float("inf"), 1), (0, 1, 0), 60.0, 64, 64),
               R.camera((0, 0, 0), (0, 5, 0), (0, 1, 0), 60.0, 64, 64),  # up parallel to view
               R.camera((2, 2, 2), (2, 2, 2), (0, 1, 0), 60.0, 64, 64)]  # eye == target
        for cam in bad:
            with pytest.raises(L.MesoError) as e:
                srv.submit(scenes[0], cam, L.RAY_SHADOW, sh)
            assert e.value.status == L.ERR_ARG
        assert srv.stats()["completed"] == 0
    finally:
        srv.stop()


def test_budget_binds_and_backpressure(ctx, scenes):
    sh = R.shading(W.LIGHT)
    # one 640x360 request needs ~3.2 MB of outputs; the budget holds one
    srv = R.Server(ctx, 4 << 20, max_pending=3, max_inflight=1)
    big = R.camera((0, 0, 0), (1, 1, 1), (0, 1, 0), 60.0, 2048, 2048)
    with pytest.raises(L.MesoError) as e:
        srv.submit(scenes[0], big, L.RAY_SHADOW, sh)
    assert e.value.status == L.ERR_ARG
    tickets, busy = [], 0
    for k in range(8):
        t = srv.submit(scenes[k % 3], _req((128, 200, 256)[k % 3], 640, 360, k), L.RAY_SHADOW, sh)
        if t is None:
            busy += 1
        else:
            tickets.append(t)
    assert busy >= 1, "pending queue of 3 never refused"
    for t in tickets:
        srv.wait(t, 30_000_000)
        srv.take(t, 640, 360)
    st = srv.stats()
    srv.stop()
    assert st["refused_busy"] == busy
    assert st["peak_bytes_in_use"] <= 4 << 20
    assert st["completed"] == len(tickets)


def _same(a, b):
    return all(np.array_equal(x.view(np.uint8), y.view(np.uint8)) for x, y in zip(a, b))


def test_ring_smaller_than_the_stream(ctx, scenes):
    """two ring entries for 24 requests: slots are reused only after their
    entry completed, and every result equals the plain path's"""
    sh = R.shading(W.LIGHT)
    srv = R.Server(ctx, 256 << 20, max_pending=64, max_inflight=2)
    sizes = [(256, 256), (640, 360), (97, 61), (1280, 720)]
    reqs = []
    try:
        for k in range(24):
            s = k % 3
            w, h = sizes[k % len(sizes)]
            cam = _req((128, 200, 256)[s], w, h, 500 + k)
            reqs.append((srv.submit(scenes[s], cam, L.RAY_SHADOW, sh), s, cam, w, h))
        got = []
        for tk, s, cam, w, h in reqs:
            srv.wait(tk, 30_000_000)
            got.append(srv.take(tk, w, h))
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == 24 and st["peak_bytes_in_use"] > 0
    for (tk, s, cam, w, h), res in zip(reqs, got):
        tg = R.RayTargets(ctx, w, h)
        R.raymarch(ctx, scenes[s], cam, L.RAY_SHADOW, sh, tg)
        ctx.sync()
        assert _same(res, tg.results()), (s, w, h)


def test_idle_exit_and_relaunch(ctx, scenes):
    """the persistent kernel returns after idle_exit_us with nothing to claim;
    a later request launches it again; work published while it runs is taken
    by the same launch"""
    sh = R.shading(W.LIGHT)
    srv = R.Server(ctx, 128 << 20, max_pending=64, max_inflight=16, idle_exit_us=300)
    try:
        cam = _req(256, 320, 200, 9)
        t1 = srv.submit(scenes[2], cam, L.RAY_SHADOW, sh)
        srv.wait(t1, 30_000_000)
        first = srv.take(t1, 320, 200)
        time.sleep(0.05)  # far past the idle exit
        n1 = srv.stats()["launches"]
        t2 = srv.submit(scenes[2], cam, L.RAY_SHADOW, sh)
        srv.wait(t2, 30_000_000)
        second = srv.take(t2, 320, 200)
        n2 = srv.stats()["launches"]
    finally:
        srv.stop()
    assert n1 >= 1 and n2 == n1 + 1
    assert _same(first, second)
    # a burst published back to back is served by one launch
    srv = R.Server(ctx, 256 << 20, max_pending=64, max_inflight=32, idle_exit_us=500000)
    try:
        tks = [srv.submit(scenes[k % 3], _req((128, 200, 256)[k % 3], 160, 120, 40 + k),
                          L.RAY_SHADOW, sh) for k in range(20)]
        for tk in tks:
            srv.wait(tk, 30_000_000)
            srv.take(tk, 160, 120, parts=("face",))
        st = srv.stats()
    finally:
        srv.stop()
    assert st["completed"] == 20 and st["launches"] == 1
