// hipGraph capture of meshing launches and the launch-shape autotuner.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <vector>

#include "meso_internal.h"

namespace meso {
meso_status launch_mesh_plain(const void *voxels, int32_t es, uint32_t mode, uint32_t *d_count,
                              uint32_t *d_quads, uint32_t capacity, hipStream_t stream);
}
using namespace meso;

static meso_status end_capture(meso_context *ctx, hipStream_t cs, meso_status st, meso_graph **out) {
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(cs, &g);
  if (st != MESO_OK) {
    if (g) hipGraphDestroy(g);
    hipStreamDestroy(cs);
    return st;
  }
  if (e != hipSuccess) {
    hipStreamDestroy(cs);
    return fail(MESO_ERR_RUNTIME, "graph capture failed: %s", hipGetErrorString(e));
  }
  hipGraphExec_t ex = nullptr;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  hipStreamDestroy(cs);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    return fail(MESO_ERR_RUNTIME, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  auto *mg = new meso_graph();
  mg->ctx = ctx;
  mg->graph = g;
  mg->exec = ex;
  *out = mg;
  return MESO_OK;
}

static float time_ms(meso_context *ctx, const std::function<meso_status()> &fn, int reps,
                     meso_status *st) {
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  std::vector<float> t;
  *st = fn();  // warm-up
  for (int i = 0; i < reps && *st == MESO_OK; i++) {
    hipEventRecord(a, ctx->stream);
    *st = fn();
    hipEventRecord(b, ctx->stream);
    hipEventSynchronize(b);
    float ms = 0;
    hipEventElapsedTime(&ms, a, b);
    t.push_back(ms);
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (t.empty()) return 1e30f;
  std::sort(t.begin(), t.end());
  return t[t.size() / 2];
}

extern "C" {

meso_status meso_mesh_graph_create(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                                   uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                                   uint64_t capacity, uint64_t *d_total, void *d_work,
                                   meso_graph **out) {
  MESO_CHECK(ctx && src && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  *out = nullptr;
  // validate through a dry path: meso_mesh_chunks checks everything it launches
  MESO_CHECK(src->size >= sizeof(meso_chunk_source) && src->voxels && d_counts && d_prefix &&
                 d_total && d_work && (src->elem_bytes == 1 || src->elem_bytes == 2) &&
                 src->n_chunks >= 1 && src->n_chunks <= 40000,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "invalid mesh graph arguments");
  MESO_HIP(hipSetDevice(ctx->device));
  hipStream_t cs;
  MESO_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    hipStreamDestroy(cs);
    return fail(MESO_ERR_RUNTIME, "begin capture: %s", hipGetErrorString(e));
  }
  meso_status st = meso_mesh_chunks(ctx, src, mode, d_counts, d_prefix, d_quads, capacity, d_total,
                                    d_work, (meso_stream)cs);
  return end_capture(ctx, cs, st, out);
}

meso_status meso_mesh_plain_graph_create(meso_context *ctx, const meso_chunk_source *src,
                                         uint32_t mode, uint32_t *d_counts, uint32_t *d_quads,
                                         uint32_t per_chunk_cap, meso_graph **out) {
  MESO_CHECK(ctx && src && out && d_counts && (per_chunk_cap == 0 || d_quads),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(src->layout == MESO_CHUNKS_PACKED && src->voxels &&
                 (src->elem_bytes == 1 || src->elem_bytes == 2) && src->n_chunks >= 1,
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "plain graphs take packed chunks");
  *out = nullptr;
  MESO_HIP(hipSetDevice(ctx->device));
  hipStream_t cs;
  MESO_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    hipStreamDestroy(cs);
    return fail(MESO_ERR_RUNTIME, "begin capture: %s", hipGetErrorString(e));
  }
  meso_status st = MESO_OK;
  for (int32_t c = 0; c < src->n_chunks && st == MESO_OK; c++)
    st = launch_mesh_plain((const uint8_t *)src->voxels + (uint64_t)c * 32768u * src->elem_bytes,
                           src->elem_bytes, mode, d_counts + c,
                           d_quads ? d_quads + (uint64_t)c * per_chunk_cap : nullptr, per_chunk_cap,
                           cs);
  return end_capture(ctx, cs, st, out);
}

meso_status meso_graph_launch(meso_graph *g, meso_stream stream) {
  MESO_CHECK(g, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL graph");
  MESO_HIP(hipGraphLaunch(g->exec, stream_of(g->ctx, stream)));
  return MESO_OK;
}

meso_status meso_graph_destroy(meso_graph *g) {
  if (!g) return MESO_OK;
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  delete g;
  return MESO_OK;
}

// Candidates: persistent grids of 1..4 workgroups per CU, the occupancy
// bound, and one workgroup per chunk.
meso_status meso_autotune_mesh(meso_context *ctx, const meso_chunk_source *src, uint32_t mode,
                               uint32_t *d_counts, uint32_t *d_prefix, uint32_t *d_quads,
                               uint64_t capacity, uint64_t *d_total, void *d_work,
                               meso_tune_result *out) {
  MESO_CHECK(ctx && src && out, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  const int cus = ctx->props.multiProcessorCount;
  std::vector<int> grids = {cus, 2 * cus, 3 * cus, 4 * cus, 0, src->n_chunks};
  meso_tune_result best{};
  best.size = sizeof(best);
  best.best_ms = 1e30f;
  int tried = 0;
  for (int g : grids) {
    MeshLaunch shape{256, g};
    meso_status st = MESO_OK;
    const float ms = time_ms(
        ctx,
        [&]() {
          return launch_mesh_batched(ctx, src, mode, d_counts, d_prefix, d_quads, capacity,
                                     d_total, d_work, ctx->stream, &shape);
        },
        5, &st);
    if (st != MESO_OK) return st;
    tried++;
    if (g == 0) best.default_ms = ms;
    if (ms < best.best_ms) {
      best.best_ms = ms;
      best.grid = g;
      best.block = 256;
      best.variant = g == 0 ? 0 : 1;
    }
  }
  best.tried = tried;
  char key[96];
  snprintf(key, sizeof(key), "mesh:%d:%d:%d:%u", src->elem_bytes, src->layout, src->n_chunks, mode);
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    ctx->tune[key] = best;
  }
  *out = best;
  return MESO_OK;
}

meso_status meso_autotune_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                                   uint32_t flags, const meso_shading *shading,
                                   const meso_ray_outputs *outp, meso_tune_result *result) {
  MESO_CHECK(ctx && scene && cam && outp && result, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "NULL argument");
  flags &= ~MESO_RAY_COUNT;
  // validate once through the public entry point
  meso_status st = meso_raymarch(ctx, scene, cam, flags, shading, outp, nullptr);
  if (st != MESO_OK) return st;
  DevCamera dc;
  derive_camera(cam, &dc);
  DevShading sh{};
  sh.light[1] = 1.0f;
  if (shading && shading->size >= sizeof(meso_shading) &&
      (shading->light[0] != 0.0f || shading->light[1] != 0.0f || shading->light[2] != 0.0f))
    derive_shading(shading, &sh);
  RayOut ro{outp->hit_index, outp->index_bytes, outp->face, outp->t, outp->shade, nullptr};
  // tiled shapes, then persistent waves with 8x8, 16x4 and 32x2 packets at
  // full occupancy (grid 0) and at 24 (16) waves per CU, then waves over
  // pairs of 8x8 packets
  const int cus = ctx->props.multiProcessorCount;
  struct Shape {
    int v, tw, th, grid;
  } shapes[] = {{0, 16, 16, 0},      {1, 32, 8, 0},       {2, 8, 8, 0},
                {3, 16, 8, 0},       {4, 32, 16, 0},      {5, 8, 8, 0},
                {5, 8, 8, 24 * cus}, {5, 8, 8, 16 * cus}, {6, 16, 4, 0},
                {6, 16, 4, 24 * cus}, {7, 32, 2, 0},      {7, 32, 2, 24 * cus},
                {8, 8, 8, 0},        {8, 8, 8, 16 * cus}};
  meso_tune_result best{};
  best.size = sizeof(best);
  best.best_ms = 1e30f;
  int tried = 0;
  for (const Shape &sp : shapes) {
    RayLaunch rl{sp.v, sp.tw * sp.th, sp.tw, sp.th, sp.grid};
    const float ms = time_ms(
        ctx, [&]() { return launch_raymarch(ctx, scene, dc, flags, sh, ro, ctx->stream, &rl); }, 5,
        &st);
    if (st != MESO_OK) return st;
    tried++;
    if (sp.v == 0) best.default_ms = ms;
    if (ms < best.best_ms) {
      best.best_ms = ms;
      best.variant = sp.v;
      best.block = sp.v >= 5 ? 64 : sp.tw * sp.th;
      best.grid = sp.grid;
      best.tile_w = sp.tw;
      best.tile_h = sp.th;
    }
  }
  best.tried = tried;
  char key[128];
  snprintf(key, sizeof(key), "ray:%d:%d:%dx%d:%u", scene->kind, scene->n, cam->width, cam->height,
           flags);
  {
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    ctx->tune[key] = best;
  }
  *result = best;
  return MESO_OK;
}

meso_status meso_autotune_pin_raymarch(meso_context *ctx, meso_scene *scene, const meso_camera *cam,
                                       uint32_t flags, const meso_tune_result *shape) {
  MESO_CHECK(ctx && scene && cam && shape, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  MESO_CHECK(cam->size >= sizeof(meso_camera) && shape->size >= sizeof(meso_tune_result),
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "struct size too small");
  static const int tiles[10][2] = {{16, 16}, {32, 8}, {8, 8}, {16, 8}, {32, 16},
                                   {8, 8},   {16, 4}, {32, 2}, {8, 8},  {16, 4}};
  meso_tune_result r = *shape;
  MESO_CHECK(r.variant >= 0 && r.variant <= 9, MESO_ERR_ARGUMENT_OUT_OF_RANGE,
             "unknown raymarch variant %d", r.variant);
  MESO_CHECK(r.tile_w == tiles[r.variant][0] && r.tile_h == tiles[r.variant][1],
             MESO_ERR_ARGUMENT_OUT_OF_RANGE, "variant %d has %dx%d %s", r.variant,
             tiles[r.variant][0], tiles[r.variant][1], r.variant <= 4 ? "tiles" : "packets");
  if (r.variant <= 4) {
    r.block = r.tile_w * r.tile_h;
    r.grid = 0;
  } else {
    MESO_CHECK(r.grid >= 0 && r.grid <= 32 * ctx->props.multiProcessorCount,
               MESO_ERR_ARGUMENT_OUT_OF_RANGE, "persistent grid %d outside 0..%d", r.grid,
               32 * ctx->props.multiProcessorCount);
    r.block = 64;
  }
  char key[128];
  snprintf(key, sizeof(key), "ray:%d:%d:%dx%d:%u", scene->kind, scene->n, cam->width, cam->height,
           flags & ~MESO_RAY_COUNT);
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  ctx->tune[key] = r;
  return MESO_OK;
}

// one line per entry: key variant block grid tile_w tile_h best_ms default_ms
meso_status meso_autotune_save(meso_context *ctx, const char *path) {
  MESO_CHECK(ctx && path, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  std::ofstream f(path);
  MESO_CHECK(f.good(), MESO_ERR_RUNTIME, "cannot write %s", path);
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  f << "# device " << ctx->props.gcnArchName << "\n";
  for (const auto &kv : ctx->tune)
    f << kv.first << " " << kv.second.variant << " " << kv.second.block << " " << kv.second.grid
      << " " << kv.second.tile_w << " " << kv.second.tile_h << " " << kv.second.best_ms << " "
      << kv.second.default_ms << "\n";
  return MESO_OK;
}

meso_status meso_autotune_load(meso_context *ctx, const char *path) {
  MESO_CHECK(ctx && path, MESO_ERR_ARGUMENT_OUT_OF_RANGE, "NULL argument");
  std::ifstream f(path);
  if (!f.good()) return MESO_OK;
  std::string line;
  std::lock_guard<std::mutex> lk(ctx->tune_mu);
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::string key;
    meso_tune_result r{};
    r.size = sizeof(r);
    if (!(is >> key >> r.variant >> r.block >> r.grid >> r.tile_w >> r.tile_h >> r.best_ms >>
          r.default_ms))
      continue;
    if (r.variant < 0 || r.variant > 9 || r.grid < 0 ||
        r.grid > 32 * ctx->props.multiProcessorCount)
      continue;
    ctx->tune[key] = r;
  }
  return MESO_OK;
}

}  // extern "C"
