"""Serving workload: raymarch requests over 16 HBM-resident terrain/cave scenes.

run() measures, on the same seeded request stream:
  - plain path: one meso_raymarch per request, first come first served, each
    request synchronised before the next starts;
  - served path: the continuous-batching server (meso_serve_*), admission by
    an HBM budget for outputs.
In both paths the client reads the face ids of each result (1 byte/pixel).
Capacity (Mrays/s) is taken with every request queued at t=0; latency (p50/p99,
arrival -> result ready on the device) at a fixed Poisson QPS, by default 60%
of the plain path's measured capacity.  Rays = primary rays + shadow rays,
counted once per request in an untimed pre-pass.
"""
import sys
import threading
import time

import numpy as np

from . import _lib as L
from . import runtime as R
from . import workloads as W

LIT_FACES = (0, 2, 4)  # +Z, +Y, +X face the light (0.4, 1, 0.3)


def make_scenes(ctx, edges, seed=4):
    scenes = [ctx.scene_generate(L.GEN_TERRAIN_CAVES, L.SCENE_BRICKMAP_U16, e, seed + i)
              for i, e in enumerate(edges)]
    ctx.sync()
    return scenes


def _cam(r):
    return R.camera(r["eye"], r["target"], r["up"], r["vfov"], r["width"], r["height"])


def _wait_until(t):
    """Sleep until shortly before t, then spin: the waiting thread must not
    hold the interpreter lock while the result consumer runs."""
    while True:
        left = t - time.perf_counter()
        if left <= 0:
            return
        if left > 3e-4:
            time.sleep(left - 2e-4)


def count_rays(ctx, scenes, reqs, sh):
    wmax = max(r["width"] * r["height"] for r in reqs)
    tg = R.RayTargets(ctx, wmax, 1)
    out = []
    for r in reqs:
        R.raymarch(ctx, scenes[r["scene"]], _cam(r), L.RAY_SHADOW, sh, tg)
        ctx.sync()
        face = tg.face.download(np.uint8, r["width"] * r["height"])
        bc = np.bincount(face, minlength=256)
        out.append(int(face.size + bc[list(LIT_FACES)].sum()))
    return out


def plain_run(ctx, scenes, reqs, sh, timed_arrivals):
    wmax = max(r["width"] * r["height"] for r in reqs)
    tg = R.RayTargets(ctx, wmax, 1)
    lat = []
    start = time.perf_counter()
    for r in reqs:
        arrival = start + (r["arrival"] if timed_arrivals else 0.0)
        _wait_until(arrival)
        R.raymarch(ctx, scenes[r["scene"]], _cam(r), L.RAY_SHADOW, sh, tg)
        ctx.sync()
        lat.append(time.perf_counter() - arrival)
        tg.face.download(np.uint8, r["width"] * r["height"])
    return np.array(lat), time.perf_counter() - start


def served_run(ctx, scenes, reqs, sh, budget_bytes, timed_arrivals, max_inflight=64,
               workgroups_per_cu=0):
    srv = R.Server(ctx, budget_bytes, max_pending=len(reqs) + 8, max_inflight=max_inflight,
                   workgroups_per_cu=workgroups_per_cu)
    lat = [None] * len(reqs)
    tickets = [None] * len(reqs)
    lag = [0.0] * len(reqs)
    ready = threading.Semaphore(0)
    errors = []

    def consumer():
        try:
            for i in range(len(reqs)):
                ready.acquire()
                r = reqs[i]
                lt = srv.wait(tickets[i], 60_000_000)
                lat[i] = lt * 1e-6 + lag[i]
                srv.take(tickets[i], r["width"], r["height"], 0, parts=("face",))
        except Exception as e:  # surfaced after join
            errors.append(e)

    th = threading.Thread(target=consumer)
    th.start()
    start = time.perf_counter()
    try:
        for i, r in enumerate(reqs):
            arrival = start + (r["arrival"] if timed_arrivals else 0.0)
            _wait_until(arrival)
            t = srv.submit(scenes[r["scene"]], _cam(r), L.RAY_SHADOW, sh)
            lag[i] = time.perf_counter() - arrival
            if t is None:
                raise RuntimeError("server refused a request (pending queue full)")
            tickets[i] = t
            ready.release()
    finally:
        th.join()
    makespan = time.perf_counter() - start
    stats = srv.stats()
    srv.stop()
    if errors:
        raise errors[0]
    return np.array(lat), makespan, stats


def write_requests(path, edges, reqs, seed=4, generator=L.GEN_TERRAIN_CAVES,
                   layout=L.SCENE_BRICKMAP_U16):
    """The request stream as the native load generator (meso_serve_bench) reads
    it: scenes (edge, generator, layout, seed) then requests (arrival, scene,
    size, camera)."""
    import struct
    with open(path, "wb") as f:
        f.write(b"MESOREQS")
        f.write(struct.pack("<Iii", 1, len(edges), len(reqs)))
        for i, e in enumerate(edges):
            f.write(struct.pack("<iiiQ", e, generator, layout, seed + i))
        for r in reqs:
            f.write(struct.pack("<diii", r["arrival"], r["scene"], r["width"], r["height"]))
            f.write(struct.pack("<10f", *r["eye"], *r["target"], *r["up"], r["vfov"]))


def run_native(n_requests, qps, budget_bytes=512 << 20, seed=4, log=print, timeout=300,
               capacity=False):
    """Replays the seeded request stream with the native load generator (no
    interpreter in the timed loop): plain and served latency at `qps`, or
    with every request queued at t=0 (capacity)."""
    import json
    import os
    import subprocess
    import tempfile
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "meso_serve_bench")
    if not os.path.exists(exe):
        raise RuntimeError("meso_serve_bench is not built (make -C mesoengine_amd/engine)")
    edges = W.serve_scenes(seed)
    reqs = W.serve_requests(edges, n_requests, qps, seed)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "requests.bin")
        write_requests(path, edges, reqs, seed)
        cmd = [exe, path, "--budget-mb", str(budget_bytes >> 20)] + (["--capacity"] if capacity
                                                                      else [])
        p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError("meso_serve_bench failed: " + p.stderr[-2000:])
    out = json.loads(p.stdout.strip().splitlines()[-1])
    log("native serving (%s, %d requests): served p50 %.2f p99 %.2f ms, %.0f Mrays/s; plain p50 "
        "%.2f p99 %.2f ms, %.0f Mrays/s" % (
            "capacity" if capacity else "%.0f QPS" % qps, n_requests, out["served"]["p50_ms"],
            out["served"]["p99_ms"], out["served"]["mrays"], out["plain"]["p50_ms"],
            out["plain"]["p99_ms"], out["plain"]["mrays"]))
    return out


def _pct(a, q):
    return float(np.percentile(a, q) * 1e3)


def _queue_stats(st):
    n = max(int(st["admitted"]), 1)
    return dict(queue_ms_mean=round(st["queue_us_sum"] / n * 1e-3, 4),
                queue_ms_max=round(st["queue_us_max"] * 1e-3, 3),
                queued_over_1ms=int(st["queued_over_1ms"]),
                admission_blocked_passes=int(st["admission_blocked_passes"]),
                refused_busy=int(st["refused_busy"]),
                peak_budget_mb=round(st["peak_bytes_in_use"] / 2**20, 1),
                launches=int(st["launches"]))


def run(ctx, n_requests=400, qps=0.0, budget_bytes=512 << 20, seed=4, log=print,
        qps_budget_bytes=32 << 20):
    """Capacity (every request queued at t=0, budget `budget_bytes`) of the
    plain and served paths, then latency at a fixed Poisson QPS with a budget
    of `qps_budget_bytes`: small enough that it binds (a 1920x1080 request
    alone holds 27 MB), so requests wait for admission while earlier results
    are still held; the queueing that causes is reported."""
    sys.setswitchinterval(1e-4)
    edges = W.serve_scenes(seed)
    t = time.time()
    scenes = make_scenes(ctx, edges, seed)
    mb = sum(s.info().device_bytes for s in scenes) / 1e6
    log("serving: %d scenes (edges %s), %.0f MB resident, built in %.1fs" % (
        len(scenes), edges, mb, time.time() - t))
    sh = R.shading(W.LIGHT)
    # capacity: everything queued at t=0
    cap_reqs = W.serve_requests(edges, min(n_requests, 200), 1.0, seed)
    cap_rays = sum(count_rays(ctx, scenes, cap_reqs, sh))
    served_run(ctx, scenes, cap_reqs[:16], sh, budget_bytes, False)  # warm the server path
    _, mk_p = plain_run(ctx, scenes, cap_reqs, sh, False)
    _, mk_s, st_s = served_run(ctx, scenes, cap_reqs, sh, budget_bytes, False)
    plain_mrays = cap_rays / mk_p / 1e6
    served_mrays = cap_rays / mk_s / 1e6
    plain_rps = len(cap_reqs) / mk_p
    if qps <= 0:
        qps = 0.6 * plain_rps
    log("serving capacity: plain %.0f Mrays/s (%.0f req/s, %.3fs), served %.0f Mrays/s (%.3fs, "
        "device busy %.3fs, %d launches); QPS %.0f" % (
            plain_mrays, plain_rps, mk_p, served_mrays, mk_s, st_s["busy_ms"] / 1e3,
            st_s["launches"], qps))
    reqs = W.serve_requests(edges, n_requests, qps, seed)
    rays_q = sum(count_rays(ctx, scenes, reqs, sh))
    lp, _ = plain_run(ctx, scenes, reqs, sh, True)
    ls, mk_q, st_q = served_run(ctx, scenes, reqs, sh, qps_budget_bytes, True)
    out = dict(qps=round(qps, 1), requests=n_requests, scenes=len(scenes), scene_edges=edges,
               resident_mb=round(mb, 1), capacity_budget_mb=budget_bytes >> 20,
               capacity_mrays_served=round(served_mrays, 1),
               capacity_mrays_plain=round(plain_mrays, 1),
               capacity_ratio=round(served_mrays / plain_mrays, 3),
               capacity_stats=_queue_stats(st_s),
               budget_mb=qps_budget_bytes >> 20,
               p50_ms_served=round(_pct(ls, 50), 3), p99_ms_served=round(_pct(ls, 99), 3),
               p50_ms_plain=round(_pct(lp, 50), 3), p99_ms_plain=round(_pct(lp, 99), 3),
               mrays_at_qps=round(rays_q / mk_q / 1e6, 1))
    out.update(_queue_stats(st_q))
    log("serving at %.0f QPS, budget %d MB: served p50 %.2f p99 %.2f ms (admission wait mean %.3f "
        "max %.3f ms, %d over 1 ms, peak %.1f MB held); plain p50 %.2f p99 %.2f ms" % (
            qps, qps_budget_bytes >> 20, out["p50_ms_served"], out["p99_ms_served"],
            out["queue_ms_mean"], out["queue_ms_max"], out["queued_over_1ms"],
            out["peak_budget_mb"], out["p50_ms_plain"], out["p99_ms_plain"]))
    for s in scenes:
        s.destroy()
    return out
