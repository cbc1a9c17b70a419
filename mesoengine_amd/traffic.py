"""HBM traffic of the hot kernels from rocprofv3 counters.

Two steps:

  run    the workload profiled under `rocprofv3 --pmc FETCH_SIZE` and again
         under `--pmc WRITE_SIZE`: three device-to-device copies of a known
         size (the calibration), three 1920x1080 primary + shadow frames of
         the 512^3 heightfield brick map in the autotuned launch shape (after
         the tuner's own candidate launches), two batched greedy and two
         batched culled meshings of the 4096-chunk batch, and one 3840x2160
         frame (primary + shadow + 4 AO) of the 2048^3 two-level caves;
  parse  reads both counter CSVs and writes a JSON summary: per kernel the
         counter's KiB per dispatch, scaled by the factor that maps the
         copy's reading to the bytes it moved (read = written = copy size).

Usage (rocprofv3 run from /tmp, the program straight after `--`):
  rocprofv3 --pmc FETCH_SIZE -d OUT/fetch -o run --output-format csv -- \
      python3 REPO/mesoengine_amd/traffic.py run
  rocprofv3 --pmc WRITE_SIZE -d OUT/write -o run --output-format csv -- \
      python3 REPO/mesoengine_amd/traffic.py run
  python3 REPO/mesoengine_amd/traffic.py parse OUT/fetch OUT/write REPO/profiles/traffic.json

Other counter passes of the same workload (L2 hits, occupancy, VALU, LDS) are
summarised per kernel with
  python3 REPO/mesoengine_amd/traffic.py parse-counters OUT.json DIR [DIR ...]
"""
import csv
import glob
import json
import os
import sys

COPY_BYTES = 1 << 30

# kernel-name prefix -> key in the summary
KERNELS = {
    "__amd_rocclr_copyBuffer": "copy_calibration",
    "void meso::raymarch_persistent_kernel<2, false, 8>": "raymarch_brickmap_terrain512_p8x8",
    "void meso::raymarch_persistent_kernel<2, false, 16>": "raymarch_brickmap_terrain512_p16x4",
    "void meso::raymarch_persistent_kernel<2, false, 32>": "raymarch_brickmap_terrain512_p32x2",
    "void meso::raymarch_kernel<2, false,": "raymarch_brickmap_terrain512_tiled",
    "void meso::raymarch_pair_kernel<2, false, 8>": "raymarch_brickmap_terrain512_pair8x8",
    "void meso::raymarch_pair_kernel<2, false, 16>": "raymarch_brickmap_terrain512_pair16x4",
    "void meso::raymarch_persistent_kernel<3, false, 8>": "raymarch_two_level_caves2048_p8x8",
    "void meso::raymarch_persistent_kernel<3, false, 16>": "raymarch_two_level_caves2048_p16x4",
    "void meso::raymarch_persistent_kernel<3, false, 32>": "raymarch_two_level_caves2048_p32x2",
    "void meso::raymarch_kernel<3, false,": "raymarch_two_level_caves2048_tiled",
    "void meso::raymarch_pair_kernel<3, false, 8>": "raymarch_two_level_caves2048_pair8x8",
    "void meso::raymarch_pair_kernel<3, false, 16>": "raymarch_two_level_caves2048_pair16x4",
    "void meso::mesh_batched_kernel<2, true, 0>": "mesh_batched_greedy_4096",
    "void meso::mesh_batched_kernel<2, false, 0>": "mesh_batched_culled_4096",
}


def _tune(ctx, L, sc, cam, flags, sh, tg):
    """the launch shape bench.py times: the autotuner's pick (its candidate
    launches are profiled too and land in the per-kernel averages)"""
    import ctypes as C
    t = L.TuneResult()
    t.size = C.sizeof(t)
    L.check(ctx.lib.meso_autotune_raymarch(ctx.h, sc.h, L.byref(cam), flags, L.byref(sh),
                                           L.byref(tg.out), L.byref(t)))
    print("traffic: tuned variant %d grid %d" % (t.variant, t.grid), flush=True)


def run():
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from mesoengine_amd import _lib as L
    from mesoengine_amd import runtime as R
    from mesoengine_amd import workloads as W

    ctx = R.Context(0)
    print("traffic: copy calibration", flush=True)
    a, b = ctx.buffer(COPY_BYTES), ctx.buffer(COPY_BYTES)
    a.zero()
    for _ in range(3):
        L.check(ctx.lib.meso_memcpy_d2d(ctx.h, b.ptr, a.ptr, COPY_BYTES, None))
    ctx.sync()
    a.free()
    b.free()

    print("traffic: 512^3 heightfield frames", flush=True)
    cfg = W.TERRAIN
    sc = ctx.scene_generate(L.GEN_HEIGHTFIELD, L.SCENE_BRICKMAP_U16, cfg["n"], cfg["seed"])
    cp = W.terrain_camera()
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], cfg["width"], cfg["height"])
    tg = R.RayTargets(ctx, cfg["width"], cfg["height"])
    _tune(ctx, L, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    for _ in range(3):
        R.raymarch(ctx, sc, cam, L.RAY_SHADOW, R.shading(W.LIGHT), tg)
    ctx.sync()
    sc.destroy()
    del tg

    print("traffic: 4096-chunk batched meshing", flush=True)
    n_chunks = W.CHUNKS["n_chunks"]
    vox = ctx.buffer(n_chunks * 32768 * 2)
    L.check(ctx.lib.meso_generate_chunks(ctx.h, W.CHUNKS["seed"], n_chunks, vox.ptr, None))
    src = R.chunk_source(vox, 2, n_chunks)
    mt = R.MeshTargets(ctx, n_chunks, 40 << 20)
    for mode in (L.MESH_GREEDY, L.MESH_GREEDY, L.MESH_CULLED, L.MESH_CULLED):
        R.mesh_chunks(ctx, src, mode, mt)
    ctx.sync()
    vox.free()
    del mt

    print("traffic: 2048^3 caves frame", flush=True)
    cfg = W.CAVES
    sc = ctx.scene_generate(L.GEN_CAVES, L.SCENE_TWO_LEVEL_U16, cfg["n"], cfg["seed"])
    cp = W.caves_camera()
    cam = R.camera(cp["eye"], cp["target"], cp["up"], cp["vfov"], cfg["width"], cfg["height"])
    tg = R.RayTargets(ctx, cfg["width"], cfg["height"], index_bytes=8)
    sh = R.shading(W.LIGHT, 5, cfg["ao_samples"], cfg["ao_radius"])
    _tune(ctx, L, sc, cam, L.RAY_SHADOW | L.RAY_AO, sh, tg)
    R.raymarch(ctx, sc, cam, L.RAY_SHADOW | L.RAY_AO, sh, tg)
    ctx.sync()
    sc.destroy()
    ctx.close()
    print("traffic: done", flush=True)
    return 0


def _read(folder, counter):
    """kernel key -> list of per-dispatch counter values (KiB).  Of the copy
    kernel only the dispatches with the largest grid count: the calibration
    copies (the library's small uploads use the same kernel)."""
    files = glob.glob(os.path.join(folder, "**", "*counter_collection*.csv"), recursive=True)
    if not files:
        raise SystemExit("no counter_collection CSV under %s" % folder)
    rows = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = row.get("Kernel_Name", "")
                for prefix, key in KERNELS.items():
                    if name.startswith(prefix):
                        rows.setdefault(key, []).append(
                            (int(row["Grid_Size"]), float(row["Counter_Value"])))
    out = {k: [v for _, v in r] for k, r in rows.items()}
    if "copy_calibration" in rows:
        g = max(g for g, _ in rows["copy_calibration"])
        out["copy_calibration"] = [v for gs, v in rows["copy_calibration"] if gs == g]
    return out


def parse(fetch_dir, write_dir, out_path):
    fetch = _read(fetch_dir, "FETCH_SIZE")
    write = _read(write_dir, "WRITE_SIZE")
    if "copy_calibration" not in fetch or "copy_calibration" not in write:
        raise SystemExit("copy kernel missing from the counter CSVs")
    kib = COPY_BYTES / 1024.0
    cf = kib / (sum(fetch["copy_calibration"]) / len(fetch["copy_calibration"]))
    cw = kib / (sum(write["copy_calibration"]) / len(write["copy_calibration"]))
    res = {"calibration": {"copy_bytes": COPY_BYTES, "fetch_factor": round(cf, 4),
                           "write_factor": round(cw, 4)},
           "kernels": {}}
    for key in KERNELS.values():
        if key == "copy_calibration" or key not in fetch or key not in write:
            continue
        f = sum(fetch[key]) / len(fetch[key])
        w = sum(write[key]) / len(write[key])
        res["kernels"][key] = {
            "dispatches": len(fetch[key]),
            "fetch_kib_raw": round(f, 1), "write_kib_raw": round(w, 1),
            "fetch_bytes": int(f * 1024 * cf), "write_bytes": int(w * 1024 * cw),
            "hbm_bytes": int(f * 1024 * cf + w * 1024 * cw)}
    with open(out_path, "w") as fo:
        json.dump(res, fo, indent=1)
        fo.write("\n")
    print(json.dumps(res, indent=1))
    return 0


CU_NUM = 256
XCC_NUM = 8  # GRBM_GUI_ACTIVE comes summed over the 8 XCDs


def _read_all(folder):
    """kernel key -> counter -> per-dispatch values, for every counter in the CSVs"""
    files = glob.glob(os.path.join(folder, "**", "*counter_collection*.csv"), recursive=True)
    out = {}
    for fn in files:
        with open(fn, newline="") as f:
            for row in csv.DictReader(f):
                name = row.get("Kernel_Name", "")
                for prefix, key in KERNELS.items():
                    if key != "copy_calibration" and name.startswith(prefix):
                        out.setdefault(key, {}).setdefault(row["Counter_Name"], []).append(
                            float(row["Counter_Value"]))
    return out


def parse_counters(out_path, folders):
    """Per-kernel averages of the counters of several passes, plus the derived
    rates: L2 hit rate, L1 (TCP) hit rate, mean resident waves per CU (of
    32), VALU busy (share of SIMD cycles issuing VALU work; the SQ counters
    are in quad-cycles), VALU lane utilisation, LDS bank-conflict ratio."""
    merged = {}
    for folder in folders:
        for key, counters in _read_all(folder).items():
            for c, vals in counters.items():
                merged.setdefault(key, {})["%s" % c] = sum(vals) / len(vals)
    res = {}
    for key, c in merged.items():
        d = {k: round(v, 1) for k, v in sorted(c.items())}
        g = c.get("GRBM_GUI_ACTIVE", 0.0) / XCC_NUM  # the kernel's active cycles
        if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c:
            d["l2_hit_rate"] = round(c["TCC_HIT_sum"] / max(1.0, c["TCC_HIT_sum"] + c["TCC_MISS_sum"]), 4)
        if g and "SQ_WAVE_CYCLES" in c:
            d["mean_waves_per_cu"] = round(4 * c["SQ_WAVE_CYCLES"] / g / CU_NUM, 2)
        if g and "SQ_ACTIVE_INST_VALU" in c:
            d["valu_busy_pct"] = round(100 * c["SQ_ACTIVE_INST_VALU"] / CU_NUM / g, 1)
        if "SQ_THREAD_CYCLES_VALU" in c and c.get("SQ_ACTIVE_INST_VALU"):
            d["valu_lane_util_pct"] = round(100 * c["SQ_THREAD_CYCLES_VALU"] / (c["SQ_ACTIVE_INST_VALU"] * 64), 1)
        if c.get("TCP_TOTAL_CACHE_ACCESSES_sum") and "TCP_TCC_READ_REQ_sum" in c:
            d["l1_hit_rate"] = round(1 - c["TCP_TCC_READ_REQ_sum"] / c["TCP_TOTAL_CACHE_ACCESSES_sum"], 4)
        if "SQ_LDS_BANK_CONFLICT" in c and c.get("SQ_LDS_IDX_ACTIVE"):
            d["lds_bank_conflict_ratio"] = round(
                c["SQ_LDS_BANK_CONFLICT"] / max(1.0, c["SQ_LDS_IDX_ACTIVE"] - c["SQ_LDS_BANK_CONFLICT"]), 4)
        res[key] = d
    with open(out_path, "w") as fo:
        json.dump(res, fo, indent=1)
        fo.write("\n")
    print(json.dumps(res, indent=1))
    return 0


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "run":
        sys.exit(run())
    if len(sys.argv) == 5 and sys.argv[1] == "parse":
        sys.exit(parse(sys.argv[2], sys.argv[3], sys.argv[4]))
    if len(sys.argv) >= 4 and sys.argv[1] == "parse-counters":
        sys.exit(parse_counters(sys.argv[2], sys.argv[3:]))
    print(__doc__)
    sys.exit(2)
