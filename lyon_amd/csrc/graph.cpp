// HIP graph forms of the batched stroke and fill, and the launch-shape autotuner.
//
// lyg_stroke_batch_graph captures one batched stroke into a HIP graph bound to the caller's
// device buffers: the default pipeline (plan, scans, cut, heads, item emit, check, scans,
// compaction, second walk of the paths the check sends back), or — with splitting off or
// staging off — counter reset, count kernel, two offset scans and the emit kernel.  A
// launch replays the whole pipeline with no host synchronisation inside; the emit kernel
// checks the output capacity on the device (a path that does not fit is skipped and
// reported), so the host reads the totals and the status only when it wants them.
// Counters are cleared by a fill kernel inside the graph (no memset nodes).
// lyg_fill_batch_graph (fill_api.cpp) captures the batched fill's first sweep round.
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
hipError_t lyg_launch_stroke_plan(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                  uint32_t split_windows, float room_scale, uint32_t* extra,
                                  uint32_t* rv, uint32_t* ri, int grid, hipStream_t s);
hipError_t lyg_launch_stroke_cut(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                 uint32_t split_windows, float room_scale, const uint32_t* extra,
                                 const uint32_t* extra_off, const uint32_t* rv,
                                 const uint32_t* ri, const uint32_t* v_off, const uint32_t* i_off,
                                 uint32_t cap_extra, uint32_t* err, int grid, hipStream_t s);
hipError_t lyg_launch_stroke_heads(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                   hipStream_t s);
hipError_t lyg_launch_stroke_items_verify(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                          hipStream_t s);
hipError_t lyg_launch_stroke_items_compact(const lyg::StrokeCompact* C,
                                           const lyg::StrokeItem* items,
                                           const lyg::StrokeHeads* heads, uint32_t n_items,
                                           int grid, hipStream_t s);
}

struct lyg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    lyg_ctx* ctx = nullptr;
    uint64_t vertex_capacity = 0, index_capacity = 0;
    // count + emit: [0] count queue, [1] emit queue, [4] overflow, [5] emit err.
    // split walk: [1] item queue, [2] second-walk queue, [3] second-walk err, [4] overflow,
    // [5] item emit err, [6] replay err, [7] paths walked again
    uint32_t* counters = nullptr;
    uint64_t* totals = nullptr;    // count + emit: [0] v, [1] i; split: [0] staging v,
                                   // [1] staging i, [2] extra parts, [3] v, [4] i
    DevBuf counts_v, counts_i;
    // split walk (the staged pipeline of lyg_stroke_batch)
    bool split = false;
    uint32_t cap_extra = 0;
    DevBuf items, heads, spec, end, parts, parts_off, rv, ri, voff, ioff, stage_v, stage_i,
        fallback;
    DevBuf scan_tiles;  // the graph's own: a later call can reallocate the context's
    lygi::FillGraph* fill = nullptr;  // a captured fill (lyg_fill_batch_graph)
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

// The default pipeline (long paths split into parts, staged emit, compaction) captured on
// fixed buffers: sized once by running the plan eagerly (capacities 1.25x that plus
// slack), then every kernel captured with the counts they need left on the device; the
// paths the check sends back are walked again by a kernel that reads their number from the
// device.  A replay whose inputs need more items, staging or output than captured reports
// LYG_ERR_CAPACITY.
bool capture_stroke_items(lyg_ctx* c, lyg_graph* g, const lyg_batch_in* in,
                          const lyg_stroke_options* opts, uint32_t n_opts,
                          const lyg_batch_out* out) {
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
    if (!ensure(c, g->heads, sizeof(lyg::StrokeHeads) * (size_t)n) ||
        !ensure(c, g->parts, 4ull * n) || !ensure(c, g->parts_off, 4ull * (n + 1)) ||
        !ensure(c, g->rv, 4ull * n) || !ensure(c, g->ri, 4ull * n) ||
        !ensure(c, g->voff, 4ull * (n + 1)) || !ensure(c, g->ioff, 4ull * (n + 1)) ||
        !ensure(c, g->fallback, 4ull * n + 4))
        return false;
    L.heads = as<lyg::StrokeHeads>(g->heads);
    uint32_t* parts = as<uint32_t>(g->parts);
    uint32_t* parts_off = as<uint32_t>(g->parts_off);
    uint32_t* rv = as<uint32_t>(g->rv);
    uint32_t* ri = as<uint32_t>(g->ri);
    uint32_t* voff = as<uint32_t>(g->voff);
    uint32_t* ioff = as<uint32_t>(g->ioff);
    const int pgrid = (int)std::min<uint32_t>(n, (uint32_t)std::max(1, c->cu_count) * 32u);
    const uint32_t W = c->stroke_split_windows;
    const float room = c->stroke_room_scale;
    auto scan = [&](const uint32_t* counts, uint32_t* offsets, uint64_t* total) {
        return check(c, scan_with(c, g->scan_tiles, counts, n, offsets, total, g->counters + 4, s),
                     "scan");
    };
    auto plan_and_scans = [&]() {
        return check(c, lyg_launch_stroke_plan(&L, n, W, room, parts, rv, ri, pgrid, s), "plan") &&
               scan(parts, parts_off, g->totals + 2) && scan(rv, voff, g->totals + 0) &&
               scan(ri, ioff, g->totals + 1);
    };
    // 1. sizes, eagerly
    uint64_t h[3] = {0, 0, 0};
    if (!check(c, lyg_launch_fill_u32(g->counters, 8, 0, s), "reset") || !plan_and_scans() ||
        !check(c, hipMemcpyAsync(h, g->totals, 24, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return false;
    const uint64_t cap_extra = h[2] + h[2] / 4 + 64;
    const uint64_t cap_sv = h[0] + h[0] / 4 + 4096, cap_si = h[1] + h[1] / 4 + 12288;
    if (cap_extra + n >= (1ull << 26) || cap_sv >= 0x7fffffffull || cap_si >= 0x7fffffffull) {
        c->last_error = "batch too large for a captured split walk";
        return false;
    }
    g->cap_extra = (uint32_t)cap_extra;
    const uint64_t cap_items = n + cap_extra;
    if (!ensure(c, g->items, sizeof(lyg::StrokeItem) * cap_items) ||
        !ensure(c, g->spec, sizeof(lyg::CarryRec) * cap_items) ||
        !ensure(c, g->end, sizeof(lyg::CarryRec) * cap_items) ||
        !ensure(c, g->stage_v, 8 * cap_sv) || !ensure(c, g->stage_i, 4 * cap_si))
        return false;
    L.items = as<lyg::StrokeItem>(g->items);
    L.item_spec = as<lyg::CarryRec>(g->spec);
    L.item_end = as<lyg::CarryRec>(g->end);
    L.n_extra_dev = reinterpret_cast<const unsigned long long*>(g->totals + 2);
    L.cap_extra = g->cap_extra;
    lyg::StrokeLaunch E = L;  // the item walk into the staging rooms
    E.out_vertices = as<float>(g->stage_v);
    E.out_indices = as<uint32_t>(g->stage_i);
    E.vertex_capacity = cap_sv;
    E.index_capacity = cap_si;
    E.vertex_offset = lyg::ITEM_ID_BIAS;
    E.local_ids = 1;
    E.work_counter = g->counters + 1;
    E.err = g->counters + 5;
    lyg::StrokeCompact C{};
    C.stage_v = as<float>(g->stage_v);
    C.stage_i = as<uint32_t>(g->stage_i);
    C.vtx_off = out->path_vertex_offsets;
    C.idx_off = out->path_index_offsets;
    C.out_vertices = out->vertices;
    C.out_indices = out->indices;
    C.vertex_offset = out->vertex_offset;
    C.local_ids = 0;
    C.n_paths = n;
    C.fallback = as<uint32_t>(g->fallback);
    C.n_fallback = g->counters + 7;
    C.n_extra_dev = L.n_extra_dev;
    C.cap_extra = g->cap_extra;
    C.out_vcap = out->vertex_capacity;
    C.out_icap = out->index_capacity;
    C.err = g->counters + 6;
    lyg::StrokeLaunch F = L;  // paths the check sends back, walked whole at their place
    F.items = nullptr;
    F.n_extra_dev = nullptr;
    F.path_list = as<uint32_t>(g->fallback);
    F.n_list_dev = g->counters + 7;
    F.vtx_off = out->path_vertex_offsets;
    F.idx_off = out->path_index_offsets;
    F.out_vertices = out->vertices;
    F.out_indices = out->indices;
    F.vertex_capacity = out->vertex_capacity;
    F.index_capacity = out->index_capacity;
    F.vertex_offset = out->vertex_offset;
    F.local_ids = 0;
    F.store_counts = 0;
    F.work_counter = g->counters + 2;
    F.err = g->counters + 3;
    const int egrid = persistent_grid(c, (uint32_t)std::min<uint64_t>(cap_items, 0xffffffffu));
    const int fgrid = std::max(1, c->cu_count) * 2;
    // 2. capture
    if (!check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"))
        return false;
    bool ok = check(c, lyg_launch_fill_u32(g->counters, 8, 0, s), "reset") && plan_and_scans();
    ok = ok && check(c, lyg_launch_stroke_cut(&L, n, W, room, parts, parts_off, rv, ri, voff, ioff,
                                              g->cap_extra, g->counters + 6, pgrid, s), "cut");
    ok = ok && check(c, lyg_launch_stroke_heads(&L, n, pgrid, s), "heads");
    ok = ok && check(c, lyg_launch_stroke(&E, 1, egrid, s), "stroke emit (items)");
    ok = ok && check(c, lyg_launch_stroke_items_verify(&L, n, pgrid, s), "verify");
    ok = ok && scan(L.path_nv, out->path_vertex_offsets, g->totals + 3);
    ok = ok && scan(L.path_ni, out->path_index_offsets, g->totals + 4);
    ok = ok && check(c, lyg_launch_stroke_items_compact(&C, L.items, L.heads, (uint32_t)cap_items,
                                                        (int)std::min<uint64_t>(cap_items, 1u << 20),
                                                        s), "compact");
    ok = ok && check(c, lyg_launch_stroke(&F, 1, fgrid, s), "stroke emit (again)");
    hipGraph_t graph = nullptr;
    const bool ended = check(c, hipStreamEndCapture(s, &graph), "end capture");
    if (!ok || !ended) {
        if (graph) (void)hipGraphDestroy(graph);
        return false;
    }
    g->graph = graph;
    g->split = true;
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
        // the scans' tile buffer is allocated before the capture, not inside it
        (in->n_paths > 32768 && !ensure(c, g->scan_tiles, lyg_scan_tiles_bytes(in->n_paths))) ||
        !(c->stroke_staged && c->stroke_split_windows > 0
              ? capture_stroke_items(c, g, in, opts, n_opts, out)
              : capture_stroke(c, g, in, opts, n_opts, out))) {
        lyg_graph_destroy(g);
        return LYG_ERR_DEVICE;
    }
    *graph = g;
    return LYG_OK;
}

int32_t lyg_fill_batch_graph(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                             uint32_t n_opts, const lyg_batch_out* out, lyg_graph** graph) {
    if (!c || !in || !opts || !out || !graph) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    if (in->n_paths == 0 || !out->path_vertex_offsets || !out->path_index_offsets)
        return LYG_ERR_INVALID_ARGUMENT;
    *graph = nullptr;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    auto* g = new lyg_graph();
    g->ctx = c;
    g->vertex_capacity = out->vertex_capacity;
    g->index_capacity = out->index_capacity;
    const int32_t st = lygi::fill_graph_capture(c, in, opts, n_opts, out, &g->fill);
    if (st != LYG_OK) {
        lyg_graph_destroy(g);
        return st;
    }
    *graph = g;
    return LYG_OK;
}

int32_t lyg_graph_launch(lyg_graph* g) {
    if (g && g->fill) return lygi::fill_graph_launch(g->fill);
    if (!g || !g->exec) return LYG_ERR_INVALID_ARGUMENT;
    return check(g->ctx, hipGraphLaunch(g->exec, g->ctx->stream), "graph launch") ? LYG_OK
                                                                                  : LYG_ERR_DEVICE;
}

int32_t lyg_graph_result(lyg_graph* g, uint64_t* n_vertices, uint64_t* n_indices) {
    if (!g) return LYG_ERR_INVALID_ARGUMENT;
    if (g->fill) return lygi::fill_graph_result(g->fill, n_vertices, n_indices);
    lyg_ctx* c = g->ctx;
    uint64_t tot[5] = {0, 0, 0, 0, 0};
    uint32_t cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!check(c, hipStreamSynchronize(c->stream), "sync") ||
        !check(c, hipMemcpy(tot, g->totals, 40, hipMemcpyDeviceToHost), "D2H") ||
        !check(c, hipMemcpy(cnt, g->counters, 32, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (g->split) {
        if (n_vertices) *n_vertices = tot[3];
        if (n_indices) *n_indices = tot[4];
        if (cnt[4]) return LYG_ERR_TOO_MANY_VERTICES;
        // staging, item slots or output arrays smaller than this replay needed
        if ((cnt[5] & 2u) || (cnt[6] & 6u) || (cnt[3] & 2u)) return LYG_ERR_CAPACITY;
        if (cnt[3]) return LYG_ERR_INTERNAL;  // the second walk wrote outside its range
        return LYG_OK;
    }
    if (n_vertices) *n_vertices = tot[0];
    if (n_indices) *n_indices = tot[1];
    if (cnt[4]) return LYG_ERR_TOO_MANY_VERTICES;
    if (cnt[5] & 2u) return LYG_ERR_CAPACITY;
    if (cnt[5]) return LYG_ERR_INTERNAL;
    return LYG_OK;
}

void lyg_graph_destroy(lyg_graph* g) {
    if (!g) return;
    lygi::fill_graph_destroy(g->fill);
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->counters) (void)hipFree(g->counters);
    if (g->totals) (void)hipFree(g->totals);
    free_buf(g->counts_v);
    free_buf(g->counts_i);
    for (DevBuf* b : {&g->items, &g->heads, &g->spec, &g->end, &g->parts, &g->parts_off, &g->rv,
                      &g->ri, &g->voff, &g->ioff, &g->stage_v, &g->stage_i, &g->fallback,
                      &g->scan_tiles})
        free_buf(*b);
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

int32_t lyg_ctx_stroke_split_stats(lyg_ctx* c, uint64_t* items, uint32_t* walked_again) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    if (items) *items = c->last_stroke_items;
    if (walked_again) *walked_again = c->last_stroke_again;
    return LYG_OK;
}

int32_t lyg_ctx_set_fill_lds_resident(lyg_ctx* c, int enabled) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    c->fill_lds_resident = enabled > 0 ? (uint32_t)enabled : 0u;
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
