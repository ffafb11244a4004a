// HIP graph form of the batched stroke and the launch-shape autotuner.
//
// lyg_stroke_batch_graph captures one batched stroke — counter reset, count kernel, two
// offset scans, emit kernel — into a HIP graph bound to the caller's device buffers.  A
// launch replays the whole pipeline with no host synchronisation inside; the emit kernel
// checks the output capacity on the device (a path that does not fit is skipped and
// reported), so the host reads the totals and the status only when it wants them.
// Counters are cleared by a fill kernel inside the graph (no memset nodes).
//
// The autotuner times the caller's batch under each candidate launch shape with HIP events
// and keeps the fastest in the context.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <vector>

#include "ctx.h"
#include "internal.h"
#include "lyg_types.h"

extern "C" {
hipError_t lyg_launch_stroke(const lyg::StrokeLaunch* L, int emit, int grid, hipStream_t stream);
hipError_t lyg_launch_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total,
                           uint32_t* overflow, hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
int lyg_fill_occupancy(uint32_t waves_per_block, uint32_t hot_bytes);
}

struct lyg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    lyg_ctx* ctx = nullptr;
    uint64_t vertex_capacity = 0, index_capacity = 0;
    uint32_t* counters = nullptr;  // [0] count queue, [1] emit queue, [4] overflow, [5] emit err
    uint64_t* totals = nullptr;
    DevBuf counts_v, counts_i;
};

namespace {

bool capture_stroke(lyg_ctx* c, lyg_graph* g, const lyg_batch_in* in,
                    const lyg_stroke_options* opts, uint32_t n_opts, const lyg_batch_out* out) {
    const uint32_t n = in->n_paths;
    hipStream_t s = c->stream;
    lyg::StrokeLaunch L{};
    L.nsub_thr = as<float>(c->nsub_thr);
    L.in.verbs = in->verbs;
    L.in.points = in->points;
    L.in.verb_off = in->verb_offsets;
    L.in.point_off = in->point_offsets;
    L.in.n_paths = n;
    L.opts = opts;
    L.opts_stride = n_opts == 1 ? 0 : 1;
    L.path_nv = as<uint32_t>(g->counts_v);
    L.path_ni = as<uint32_t>(g->counts_i);
    L.vertex_offset = out->vertex_offset;
    L.vtx_off = out->path_vertex_offsets;
    L.idx_off = out->path_index_offsets;
    L.out_vertices = out->vertices;
    L.out_indices = out->indices;
    L.vertex_capacity = out->vertex_capacity;
    L.index_capacity = out->index_capacity;
    L.err = g->counters + 5;
    const int grid = persistent_grid(c, n);
    if (!check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"))
        return false;
    bool ok = check(c, lyg_launch_fill_u32(g->counters, 8, 0, s), "reset");
    L.work_counter = g->counters + 0;
    ok = ok && check(c, lyg_launch_stroke(&L, 0, grid, s), "stroke count");
    ok = ok && check(c, lyg_launch_scan(L.path_nv, n, out->path_vertex_offsets, g->totals + 0,
                                        g->counters + 4, s), "scan");
    ok = ok && check(c, lyg_launch_scan(L.path_ni, n, out->path_index_offsets, g->totals + 1,
                                        g->counters + 4, s), "scan");
    L.work_counter = g->counters + 1;
    ok = ok && check(c, lyg_launch_stroke(&L, 1, grid, s), "stroke emit");
    hipGraph_t graph = nullptr;
    const bool ended = check(c, hipStreamEndCapture(s, &graph), "end capture");
    if (!ok || !ended) {
        if (graph) (void)hipGraphDestroy(graph);
        return false;
    }
    g->graph = graph;
    return check(c, hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0), "instantiate");
}

}  // namespace

extern "C" {

int32_t lyg_stroke_batch_graph(lyg_ctx* c, const lyg_batch_in* in, const lyg_stroke_options* opts,
                               uint32_t n_opts, const lyg_batch_out* out, lyg_graph** graph) {
    if (!c || !in || !opts || !out || !graph) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    if (in->n_paths == 0) return LYG_ERR_INVALID_ARGUMENT;
    *graph = nullptr;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    auto* g = new lyg_graph();
    g->ctx = c;
    g->vertex_capacity = out->vertex_capacity;
    g->index_capacity = out->index_capacity;
    if (!check(c, hipMalloc(&g->counters, 64), "hipMalloc") ||
        !check(c, hipMalloc(&g->totals, 64), "hipMalloc") ||
        !ensure(c, g->counts_v, 4ull * in->n_paths) || !ensure(c, g->counts_i, 4ull * in->n_paths) ||
        !capture_stroke(c, g, in, opts, n_opts, out)) {
        lyg_graph_destroy(g);
        return LYG_ERR_DEVICE;
    }
    *graph = g;
    return LYG_OK;
}

int32_t lyg_graph_launch(lyg_graph* g) {
    if (!g || !g->exec) return LYG_ERR_INVALID_ARGUMENT;
    return check(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream), "graph launch") ? LYG_OK
                                                                                  : LYG_ERR_DEVICE;
}

int32_t lyg_graph_result(lyg_graph* g, uint64_t* n_vertices, uint64_t* n_indices) {
    if (!g) return LYG_ERR_INVALID_ARGUMENT;
    lyg_ctx* c = g->ctx;
    uint64_t tot[2] = {0, 0};
    uint32_t flags[2] = {0, 0};
    if (!check(c, hipStreamSynchronize(c->stream), "sync") ||
        !check(c, hipMemcpy(tot, g->totals, 16, hipMemcpyDeviceToHost), "D2H") ||
        !check(c, hipMemcpy(flags, g->counters + 4, 8, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (n_vertices) *n_vertices = tot[0];
    if (n_indices) *n_indices = tot[1];
    if (flags[0]) return LYG_ERR_TOO_MANY_VERTICES;
    if (flags[1] & 2u) return LYG_ERR_CAPACITY;
    if (flags[1]) return LYG_ERR_INTERNAL;
    return LYG_OK;
}

void lyg_graph_destroy(lyg_graph* g) {
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->counters) (void)hipFree(g->counters);
    if (g->totals) (void)hipFree(g->totals);
    free_buf(g->counts_v);
    free_buf(g->counts_i);
    delete g;
}

// ---- autotuner -----------------------------------------------------------------------
int32_t lyg_ctx_set_tuning(lyg_ctx* c, const lyg_tuning* stroke, const lyg_tuning* fill) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    if (stroke) {
        if (stroke->blocks_per_cu < 1 || stroke->blocks_per_cu > 64) return LYG_ERR_INVALID_ARGUMENT;
        c->stroke_blocks_per_cu = stroke->blocks_per_cu;
    }
    if (fill) {
        if (fill->waves_per_block < 1 || fill->waves_per_block > 4 ||
            lyg_fill_occupancy(fill->waves_per_block, fill->lds_bytes_per_wave) < 1)
            return LYG_ERR_INVALID_ARGUMENT;
        c->fill_waves_per_block = fill->waves_per_block;
        c->fill_hot_bytes = fill->lds_bytes_per_wave;
    }
    return LYG_OK;
}

int32_t lyg_ctx_set_stroke_staging(lyg_ctx* c, int enabled, float room_scale) {
    if (!c || !(room_scale > 0.0f) || room_scale > 64.0f) return LYG_ERR_INVALID_ARGUMENT;
    c->stroke_staged = enabled != 0;
    c->stroke_room_scale = room_scale;
    return LYG_OK;
}

int32_t lyg_ctx_set_stroke_split(lyg_ctx* c, uint32_t windows_per_part) {
    if (!c || windows_per_part > (1u << 20)) return LYG_ERR_INVALID_ARGUMENT;
    c->stroke_split_windows = windows_per_part;
    return LYG_OK;
}

int32_t lyg_autotune_stroke(lyg_ctx* c, const lyg_batch_in* in, const lyg_stroke_options* opts,
                            uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    static const uint32_t cand[] = {2, 4, 6, 8, 12, 16};
    hipEvent_t e0, e1;
    if (!check(c, hipEventCreate(&e0), "event") || !check(c, hipEventCreate(&e1), "event"))
        return LYG_ERR_DEVICE;
    const uint32_t keep = c->stroke_blocks_per_cu;
    float best = 1e30f;
    uint32_t best_b = keep;
    int32_t st = LYG_OK;
    for (uint32_t b : cand) {
        c->stroke_blocks_per_cu = b;
        float ms = 0;
        for (int rep = 0; rep < 2 && st == LYG_OK; ++rep) {  // first run warms up
            (void)hipEventRecord(e0, c->stream);
            st = lyg_stroke_batch(c, in, opts, n_opts, out);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (st != LYG_OK) break;
        if (ms < best) {
            best = ms;
            best_b = b;
        }
    }
    c->stroke_blocks_per_cu = st == LYG_OK ? best_b : keep;
    // then the parts long paths are cut into (0: whole paths)
    static const uint32_t split_cand[] = {0, 16, 32, 64, 128};
    const uint32_t keep_w = c->stroke_split_windows;
    uint32_t best_w = keep_w;
    float best_split = 1e30f;
    for (uint32_t w : split_cand) {
        if (st != LYG_OK) break;
        c->stroke_split_windows = w;
        float ms = 0;
        for (int rep = 0; rep < 2 && st == LYG_OK; ++rep) {
            (void)hipEventRecord(e0, c->stream);
            st = lyg_stroke_batch(c, in, opts, n_opts, out);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        if (st == LYG_OK && ms < best_split) {
            best_split = ms;
            best_w = w;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->stroke_split_windows = st == LYG_OK ? best_w : keep_w;
    if (chosen) {
        chosen->blocks_per_cu = c->stroke_blocks_per_cu;
        chosen->waves_per_block = 4;
        chosen->lds_bytes_per_wave = 0;
        chosen->best_ms = std::min(best, best_split);
        chosen->split_windows = c->stroke_split_windows;
        chosen->edges_per_band = c->band_edges_per_band;
    }
    return st;
}

int32_t lyg_autotune_banded(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                            uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    static const uint32_t cand[] = {16, 32, 64, 128, 256};
    hipEvent_t e0, e1;
    if (!check(c, hipEventCreate(&e0), "event") || !check(c, hipEventCreate(&e1), "event"))
        return LYG_ERR_DEVICE;
    const uint32_t keep = c->band_edges_per_band;
    float best = 1e30f;
    uint32_t best_e = keep;
    int32_t st = LYG_OK;
    for (uint32_t e : cand) {
        c->band_edges_per_band = e;
        float ms = 0;
        for (int rep = 0; rep < 2; ++rep) {  // first run warms up
            (void)hipEventRecord(e0, c->stream);
            st = lyg_fill_batch_banded(c, in, opts, n_opts, out);
            (void)hipEventRecord(e1, c->stream);
            (void)hipEventSynchronize(e1);
            (void)hipEventElapsedTime(&ms, e0, e1);
        }
        // per-path failures (a path's own status) do not disqualify a band size; one
        // whose output does not fit the caller's arrays is skipped
        if (st == LYG_ERR_CAPACITY) {
            st = LYG_OK;
            continue;
        }
        if (st != LYG_OK && st != LYG_ERR_INTERNAL && st != LYG_ERR_POSITION_IS_NAN &&
            st != LYG_ERR_TOLERANCE_IS_NAN)
            break;
        st = LYG_OK;
        if (ms < best) {
            best = ms;
            best_e = e;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->band_edges_per_band = st == LYG_OK ? best_e : keep;
    if (chosen) {
        chosen->blocks_per_cu = 0;
        chosen->waves_per_block = 1;
        chosen->lds_bytes_per_wave = 0;
        chosen->best_ms = best;
        chosen->split_windows = c->stroke_split_windows;
        chosen->edges_per_band = c->band_edges_per_band;
    }
    return st;
}

int32_t lyg_autotune_fill(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                          uint32_t n_opts, lyg_batch_out* out, lyg_tuning* chosen) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    struct Shape {
        uint32_t wpb, hot;
    };
    static const Shape cand[] = {{1, 64 * 1024}, {1, 40 * 1024}, {2, 32 * 1024}, {2, 24 * 1024},
                                 {2, 16 * 1024}, {4, 16 * 1024}, {4, 8 * 1024},  {2, 0}};
    hipEvent_t e0, e1;
    if (!check(c, hipEventCreate(&e0), "event") || !check(c, hipEventCreate(&e1), "event"))
        return LYG_ERR_DEVICE;
    const uint32_t keep_w = c->fill_waves_per_block, keep_h = c->fill_hot_bytes;
    (void)lyg_fill_batch(c, in, opts, n_opts, out);  // warm-up: allocations, code loading
    float best = 1e30f;
    Shape best_s{keep_w, keep_h};
    int32_t st = LYG_OK;
    for (const Shape& sh : cand) {
        if (lyg_fill_occupancy(sh.wpb, sh.hot) < 1) continue;  // does not fit a CU
        c->fill_waves_per_block = sh.wpb;
        c->fill_hot_bytes = sh.hot;
        float ms = 0;
        (void)hipEventRecord(e0, c->stream);
        st = lyg_fill_batch(c, in, opts, n_opts, out);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        // per-path failures (a path's own status) do not disqualify a shape
        if (st != LYG_OK && st != LYG_ERR_INTERNAL && st != LYG_ERR_POSITION_IS_NAN &&
            st != LYG_ERR_TOLERANCE_IS_NAN && st != LYG_ERR_NO_SPACE)
            break;
        st = LYG_OK;
        if (ms < best) {
            best = ms;
            best_s = sh;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->fill_waves_per_block = st == LYG_OK ? best_s.wpb : keep_w;
    c->fill_hot_bytes = st == LYG_OK ? best_s.hot : keep_h;
    if (chosen) {
        chosen->blocks_per_cu = 0;
        chosen->waves_per_block = c->fill_waves_per_block;
        chosen->lds_bytes_per_wave = c->fill_hot_bytes;
        chosen->best_ms = best;
        chosen->split_windows = c->stroke_split_windows;
        chosen->edges_per_band = c->band_edges_per_band;
    }
    return st;
}

}  // extern "C"
