// Host runtime of liblyg.so: contexts, the device buffer pool and the C ABI.
//
// A context owns one HIP stream (default priority) and a small set of grow-only device
// buffers.  The plain per-path calls stage the path into device memory, run the count
// kernel, size-check the caller's arrays (two-phase protocol), run the emit kernel and
// copy the geometry back — the same kernels the batched form runs over many paths.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lyg.h"
#include "lyg_types.h"

extern "C" {
hipError_t lyg_launch_stroke(const lyg::StrokeLaunch* L, int emit, int grid, hipStream_t stream);
hipError_t lyg_launch_stroke_stage_caps(const lyg::StrokeLaunch* L, uint32_t n, float room_scale,
                                        uint32_t* cap_v, uint32_t* cap_i, hipStream_t s);
hipError_t lyg_launch_stroke_compact(const lyg::StrokeCompact* C, int grid, hipStream_t s);
hipError_t lyg_launch_stroke_plan(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                  uint32_t split_windows, float room_scale, uint32_t* extra,
                                  uint32_t* rv, uint32_t* ri, int grid, hipStream_t s);
hipError_t lyg_launch_stroke_cut(const lyg::StrokeLaunch* L, uint32_t n_paths,
                                 uint32_t split_windows, float room_scale, const uint32_t* extra,
                                 const uint32_t* extra_off, const uint32_t* rv,
                                 const uint32_t* ri, const uint32_t* v_off, const uint32_t* i_off,
                                 uint32_t cap_extra, uint32_t* err, int grid, hipStream_t s);
hipError_t lyg_launch_stroke_items_verify(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                          hipStream_t s);
hipError_t lyg_launch_stroke_heads(const lyg::StrokeLaunch* L, uint32_t n_paths, int grid,
                                   hipStream_t s);
hipError_t lyg_launch_stroke_items_compact(const lyg::StrokeCompact* C,
                                           const lyg::StrokeItem* items,
                                           const lyg::StrokeHeads* heads, uint32_t n_items,
                                           int grid, hipStream_t s);
hipError_t lyg_launch_nsub_thresholds(float* thr, hipStream_t s);
hipError_t lyg_launch_flatten_count(const lyg::FlattenLaunch* L, hipStream_t s);
hipError_t lyg_launch_flatten_emit(const lyg::FlattenLaunch* L, uint32_t total, hipStream_t s);
size_t lyg_flatten_runs_bytes(uint32_t total);
size_t lyg_flatten_poly_bytes(uint32_t n);
hipError_t lyg_launch_arc_flatten(const float* arcs, uint32_t n, float tol, uint32_t* counts,
                                  const uint32_t* offsets, float* out, hipStream_t s);
hipError_t lyg_launch_arc_to_quads(const float* arcs, uint32_t n, uint32_t* counts,
                                   const uint32_t* offsets, float* out, hipStream_t s);
hipError_t lyg_launch_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total,
                           uint32_t* overflow, hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
size_t lyg_scan_tiles_bytes(uint32_t n);
hipError_t lyg_launch_scan_tiles(const uint32_t* counts, uint32_t n, uint32_t* offsets,
                                 uint64_t* total, uint32_t* overflow, uint64_t* tiles,
                                 hipStream_t s);
hipError_t lyg_launch_copy(void* dst, const void* src, uint64_t bytes, int grid, hipStream_t s);
uint32_t lyg_noalloc_max_segments();
hipError_t lyg_noalloc_scratch_init(int device);
void lyg_noalloc_scratch(int device, uint32_t** counters, uint64_t** totals, uint32_t** counts,
                         uint64_t** tiles);
}

#include "ctx.h"
#include "internal.h"

namespace {

// Count pass + scan over the paths of `L`; fills per-path counts and offsets on the
// device and the totals on the host.
int32_t stroke_count_and_scan(lyg_ctx* c, lyg::StrokeLaunch& L, uint32_t* offs_v,
                              uint32_t* offs_i, uint64_t* total_v, uint64_t* total_i) {
    const uint32_t n = L.in.n_paths;
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, c->stream), "reset")) return LYG_ERR_DEVICE;
    L.work_counter = c->d_counters + 0;
    if (c->timing) (void)hipEventRecord(c->ev[0], c->stream);
    if (!check(c, lyg_launch_stroke(&L, 0, persistent_grid(c, n), c->stream), "stroke count"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[1], c->stream);
    if (!check(c, ctx_scan(c, L.path_nv, n, offs_v, c->d_totals + 0, c->d_counters + 4, c->stream),
               "scan"))
        return LYG_ERR_DEVICE;
    if (!check(c, ctx_scan(c, L.path_ni, n, offs_i, c->d_totals + 1, c->d_counters + 4, c->stream),
               "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, c->stream), "copy"))
        return LYG_ERR_DEVICE;
    if (!check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, c->stream),
               "copy"))
        return LYG_ERR_DEVICE;
    if (!check(c, hipStreamSynchronize(c->stream), "sync")) return LYG_ERR_DEVICE;
    *total_v = h[0];
    *total_i = h[1];
    uint32_t ovf = (uint32_t)(h[2] & 0xffffffffu);
    if (ovf) return LYG_ERR_TOO_MANY_VERTICES;
    return LYG_OK;
}

int32_t stroke_emit(lyg_ctx* c, lyg::StrokeLaunch& L) {
    if (!check(c, lyg_launch_fill_u32(c->d_counters + 1, 1, 0, c->stream), "reset"))
        return LYG_ERR_DEVICE;
    L.work_counter = c->d_counters + 1;
    L.err = c->d_counters + 5;
    if (c->timing) (void)hipEventRecord(c->ev[2], c->stream);
    if (!check(c, lyg_launch_stroke(&L, 1, persistent_grid(c, L.in.n_paths), c->stream),
               "stroke emit"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[3], c->stream);
    uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 128);
    if (!check(c, hipMemcpyAsync(h, c->d_counters + 5, 4, hipMemcpyDeviceToHost, c->stream),
               "copy"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

// After the stream has been synchronized: did the emit pass flag an inconsistency?
bool emit_failed(lyg_ctx* c) {
    const uint32_t* h = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(c->h_pinned) + 128);
    if (*h) {
        c->last_error = (*h & 2u) ? "a path did not fit the output arrays"
                                  : "emit pass wrote outside its path range (count/emit mismatch)";
        return true;
    }
    return false;
}

}  // namespace

namespace {

constexpr size_t STROKE_STAGE_BUDGET = 32ull << 30;  // staging bytes of one batch

// Staged single pass: every path is emitted once into its own staging room (sized by
// stroke_stage_caps_kernel), with path-local ids, while the exact counts are stored; the
// counts are scanned into the final offsets, the caller's arrays are allocated, and one
// copy moves each path into place.  A path that outgrew its room is emitted again
// directly into the final arrays.  Returns 1 when the batch does not fit the staging
// budget (the caller then runs the count and emit passes).
int32_t stroke_batch_staged(lyg_ctx* c, const lygi::BatchSpec& b, const lygi::OutAlloc& alloc,
                            lyg::StrokeLaunch& L, uint64_t* tv_out, uint64_t* ti_out) {
    const uint32_t n = b.in.n_paths;
    hipStream_t s = c->stream;
    if (!ensure(c, c->st_cap_v, 4ull * n) || !ensure(c, c->st_cap_i, 4ull * n) ||
        !ensure(c, c->st_off_v, 4ull * (n + 1)) || !ensure(c, c->st_off_i, 4ull * (n + 1)) ||
        !ensure(c, c->st_fallback, 4ull * n + 4))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    // 1. staging room and its offsets
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, lyg_launch_stroke_stage_caps(&L, n, c->stroke_room_scale,
                                               as<uint32_t>(c->st_cap_v), as<uint32_t>(c->st_cap_i), s),
               "stroke stage caps") ||
        !check(c, ctx_scan(c, as<uint32_t>(c->st_cap_v), n, as<uint32_t>(c->st_off_v),
                                  c->d_totals + 0, c->d_counters + 4, s),
               "scan") ||
        !check(c, ctx_scan(c, as<uint32_t>(c->st_cap_i), n, as<uint32_t>(c->st_off_i),
                                  c->d_totals + 1, c->d_counters + 4, s),
               "scan"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t sv = h[0], si = h[1];
    if ((uint32_t)h[2] != 0 || 8 * sv + 4 * si > STROKE_STAGE_BUDGET) return 1;
    if (!ensure(c, c->st_v, 8 * sv) || !ensure(c, c->st_i, 4 * si)) return 1;
    // 2. one emit pass into the staging room, storing the exact counts
    L.vtx_off = as<uint32_t>(c->st_off_v);
    L.idx_off = as<uint32_t>(c->st_off_i);
    L.out_vertices = as<float>(c->st_v);
    L.out_indices = as<uint32_t>(c->st_i);
    L.vertex_capacity = sv;
    L.index_capacity = si;
    L.vertex_offset = 0;
    L.local_ids = 1;
    L.store_counts = 1;
    L.work_counter = c->d_counters + 1;
    L.err = c->d_counters + 5;  // dropped writes are expected here: the counts decide
    if (c->timing) (void)hipEventRecord(c->ev[2], s);
    if (!check(c, lyg_launch_stroke(&L, 1, persistent_grid(c, n), s), "stroke emit (staged)"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[3], s);
    // 3. final offsets from the exact counts
    if (!check(c, lyg_launch_fill_u32(c->d_counters + 4, 1, 0, s), "reset") ||
        !check(c, ctx_scan(c, L.path_nv, n, b.vtx_off, c->d_totals + 0, c->d_counters + 4, s),
               "scan") ||
        !check(c, ctx_scan(c, L.path_ni, n, b.idx_off, c->d_totals + 1, c->d_counters + 4, s),
               "scan") ||
        !check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t tv = h[0], ti = h[1];
    *tv_out = tv;
    *ti_out = ti;
    if ((uint32_t)h[2] != 0) return LYG_ERR_TOO_MANY_VERTICES;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (!b.local_ids && tv > 0 && (uint64_t)b.vertex_offset + tv - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    int32_t st = alloc(tv, ti, &ov, &oi);
    if (st != LYG_OK) return st;
    // 4. compaction, and the list of paths that outgrew their room
    uint32_t* n_fb = as<uint32_t>(c->st_fallback) + n;
    lyg::StrokeCompact C{};
    C.path_nv = L.path_nv;
    C.path_ni = L.path_ni;
    C.stage_voff = as<uint32_t>(c->st_off_v);
    C.stage_ioff = as<uint32_t>(c->st_off_i);
    C.stage_v = as<float>(c->st_v);
    C.stage_i = as<uint32_t>(c->st_i);
    C.vtx_off = b.vtx_off;
    C.idx_off = b.idx_off;
    C.out_vertices = ov;
    C.out_indices = oi;
    C.vertex_offset = b.vertex_offset;
    C.local_ids = b.local_ids ? 1u : 0u;
    C.n_paths = n;
    C.fallback = as<uint32_t>(c->st_fallback);
    C.n_fallback = n_fb;
    uint32_t* hf = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 64);
    if (!check(c, lyg_launch_fill_u32(n_fb, 1, 0, s), "reset") ||
        !check(c, lyg_launch_stroke_compact(&C, (int)std::min<uint32_t>(n, 1u << 20), s),
               "stroke compact") ||
        !check(c, hipMemcpyAsync(hf, n_fb, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    // 5. paths that outgrew their room: emitted again at their final offsets
    if (*hf > 0) {
        L.in.n_paths = *hf;
        L.path_list = as<uint32_t>(c->st_fallback);
        L.vtx_off = b.vtx_off;
        L.idx_off = b.idx_off;
        L.out_vertices = ov;
        L.out_indices = oi;
        L.vertex_capacity = tv;
        L.index_capacity = ti;
        L.vertex_offset = b.vertex_offset;
        L.local_ids = b.local_ids ? 1u : 0u;
        L.store_counts = 0;
        L.work_counter = c->d_counters + 2;
        L.err = c->d_counters + 6;
        uint32_t* he = hf + 1;
        if (!check(c, lyg_launch_fill_u32(c->d_counters + 2, 1, 0, s), "reset") ||
            !check(c, lyg_launch_fill_u32(c->d_counters + 6, 1, 0, s), "reset") ||
            !check(c, lyg_launch_stroke(&L, 1, persistent_grid(c, *hf), s), "stroke emit (again)") ||
            !check(c, hipMemcpyAsync(he, c->d_counters + 6, 4, hipMemcpyDeviceToHost, s), "copy") ||
            !check(c, hipStreamSynchronize(s), "sync"))
            return LYG_ERR_DEVICE;
        if (*he) {
            c->last_error = "emit pass wrote outside its path range (staged/final mismatch)";
            return LYG_ERR_INTERNAL;
        }
    }
    if (trace_on()) {
        fprintf(stderr, "[lyg] stroke staged: %u paths, staging %llu v / %llu i, out %llu v / %llu i, %u emitted again\n",
                n, (unsigned long long)sv, (unsigned long long)si, (unsigned long long)tv,
                (unsigned long long)ti, *hf);
        fflush(stderr);
    }
    return LYG_OK;
}

// Staged single pass over items: every path is one item, except paths of one sub-path
// with at least 2 * stroke_split_windows windows, which are cut into parts walked by
// different waves (stroke_items_kernel).  Each item is emitted into its own staging room
// with item-local ids; a check per path confirms that each part's prologue rebuilt the
// carry the previous part ended with (stroke_items_verify_kernel: else the path is
// counted whole and emitted again whole at its final place); the compaction places every
// item in its path.  The output is the same as the per-path staged pass.  Returns 1 when
// the batch does not fit the staging budget.
int32_t stroke_batch_items(lyg_ctx* c, const lygi::BatchSpec& b, const lygi::OutAlloc& alloc,
                           lyg::StrokeLaunch& L, uint64_t* tv_out, uint64_t* ti_out) {
    const uint32_t n = b.in.n_paths;
    hipStream_t s = c->stream;
    if (!ensure(c, c->st_heads, sizeof(lyg::StrokeHeads) * (size_t)n) ||
        !ensure(c, c->st_cap_v, 4ull * n) || !ensure(c, c->st_cap_i, 4ull * n) ||
        !ensure(c, c->st_off_v, 4ull * (n + 1)) || !ensure(c, c->st_off_i, 4ull * (n + 1)) ||
        !ensure(c, c->st_parts, 4ull * n) || !ensure(c, c->st_parts_off, 4ull * (n + 1)) ||
        !ensure(c, c->st_fallback, 4ull * n + 4))
        return LYG_ERR_DEVICE;
    L.heads = as<lyg::StrokeHeads>(c->st_heads);
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    const int pgrid = (int)std::min<uint32_t>(n, (uint32_t)std::max(1, c->cu_count) * 32u);
    uint32_t* parts = as<uint32_t>(c->st_parts);
    uint32_t* parts_off = as<uint32_t>(c->st_parts_off);
    uint32_t* rv = as<uint32_t>(c->st_cap_v);
    uint32_t* ri = as<uint32_t>(c->st_cap_i);
    uint32_t* voff = as<uint32_t>(c->st_off_v);
    uint32_t* ioff = as<uint32_t>(c->st_off_i);
    // 1. the plan of every path (parts beyond the first, staging rooms), its offsets
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, lyg_launch_stroke_plan(&L, n, c->stroke_split_windows, c->stroke_room_scale,
                                         parts, rv, ri, pgrid, s),
               "stroke plan") ||
        !check(c, ctx_scan(c, parts, n, parts_off, c->d_totals + 2, c->d_counters + 4, s), "scan") ||
        !check(c, ctx_scan(c, rv, n, voff, c->d_totals + 0, c->d_counters + 4, s), "scan") ||
        !check(c, ctx_scan(c, ri, n, ioff, c->d_totals + 1, c->d_counters + 4, s), "scan") ||
        !check(c, hipMemcpyAsync(h, c->d_totals, 24, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipMemcpyAsync(h + 3, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t sv = h[0], si = h[1], n_extra = h[2];
    const uint64_t n_items = (uint64_t)n + n_extra;
    if ((uint32_t)h[3] != 0 || sv >= 0x7fffffffull || si >= 0x7fffffffull ||
        n_items >= (1ull << 26) || 8 * sv + 4 * si > STROKE_STAGE_BUDGET)
        return 1;
    if (!ensure(c, c->st_items, sizeof(lyg::StrokeItem) * n_items) ||
        !ensure(c, c->st_spec, sizeof(lyg::CarryRec) * n_items) ||
        !ensure(c, c->st_end, sizeof(lyg::CarryRec) * n_items))
        return LYG_ERR_DEVICE;
    if (!ensure(c, c->st_v, 8 * sv + 8) || !ensure(c, c->st_i, 4 * si + 4)) return 1;
    L.items = as<lyg::StrokeItem>(c->st_items);
    L.item_spec = as<lyg::CarryRec>(c->st_spec);
    L.item_end = as<lyg::CarryRec>(c->st_end);
    // the items, and the first points of split paths
    if (!check(c, lyg_launch_stroke_cut(&L, n, c->stroke_split_windows, c->stroke_room_scale,
                                        parts, parts_off, rv, ri, voff, ioff,
                                        (uint32_t)n_extra, nullptr, pgrid, s),
               "stroke cut") ||
        !check(c, lyg_launch_stroke_heads(&L, n, pgrid, s), "stroke heads"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    // 2. one emit pass over the items into their staging rooms
    L.n_items = (uint32_t)n_items;
    L.out_vertices = as<float>(c->st_v);
    L.out_indices = as<uint32_t>(c->st_i);
    L.vertex_capacity = sv;
    L.index_capacity = si;
    L.vertex_offset = lyg::ITEM_ID_BIAS;  // item-local ids count from the bias
    L.local_ids = 1;
    L.store_counts = 0;
    L.work_counter = c->d_counters + 1;
    L.err = c->d_counters + 5;  // dropped writes are expected here: the counts decide
    if (c->timing) (void)hipEventRecord(c->ev[2], s);
    if (!check(c, lyg_launch_stroke(&L, 1, persistent_grid(c, (uint32_t)std::min<uint64_t>(n_items, 0xffffffffu)), s),
               "stroke emit (items)"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[3], s);
    // 3. check the parts, place them in their paths, the paths' counts and final offsets
    if (!check(c, lyg_launch_stroke_items_verify(&L, n, pgrid, s), "stroke items verify") ||
        !check(c, lyg_launch_fill_u32(c->d_counters + 4, 1, 0, s), "reset") ||
        !check(c, ctx_scan(c, L.path_nv, n, b.vtx_off, c->d_totals + 0, c->d_counters + 4, s),
               "scan") ||
        !check(c, ctx_scan(c, L.path_ni, n, b.idx_off, c->d_totals + 1, c->d_counters + 4, s),
               "scan") ||
        !check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t tv = h[0], ti = h[1];
    *tv_out = tv;
    *ti_out = ti;
    if ((uint32_t)h[2] != 0) return LYG_ERR_TOO_MANY_VERTICES;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (!b.local_ids && tv > 0 && (uint64_t)b.vertex_offset + tv - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    int32_t st = alloc(tv, ti, &ov, &oi);
    if (st != LYG_OK) return st;
    // 4. compaction; paths marked for a whole walk are listed
    uint32_t* n_fb = as<uint32_t>(c->st_fallback) + n;
    lyg::StrokeCompact C{};
    C.stage_v = as<float>(c->st_v);
    C.stage_i = as<uint32_t>(c->st_i);
    C.vtx_off = b.vtx_off;
    C.idx_off = b.idx_off;
    C.out_vertices = ov;
    C.out_indices = oi;
    C.vertex_offset = b.vertex_offset;
    C.local_ids = b.local_ids ? 1u : 0u;
    C.n_paths = n;
    C.fallback = as<uint32_t>(c->st_fallback);
    C.n_fallback = n_fb;
    C.out_vcap = tv;
    C.out_icap = ti;
    uint32_t* hf = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 64);
    if (!check(c, lyg_launch_fill_u32(n_fb, 1, 0, s), "reset") ||
        !check(c, lyg_launch_stroke_items_compact(&C, L.items, L.heads, L.n_items,
                                                  (int)std::min<uint64_t>(n_items, 1u << 20), s),
               "stroke items compact") ||
        !check(c, hipMemcpyAsync(hf, n_fb, 4, hipMemcpyDeviceToHost, s), "copy") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    // 5. listed paths: emitted again whole at their final offsets
    if (*hf > 0) {
        lyg::StrokeLaunch F = L;
        F.items = nullptr;
        F.in.n_paths = *hf;
        F.path_list = as<uint32_t>(c->st_fallback);
        F.vtx_off = b.vtx_off;
        F.idx_off = b.idx_off;
        F.out_vertices = ov;
        F.out_indices = oi;
        F.vertex_capacity = tv;
        F.index_capacity = ti;
        F.vertex_offset = b.vertex_offset;
        F.local_ids = b.local_ids ? 1u : 0u;
        F.store_counts = 0;
        F.work_counter = c->d_counters + 2;
        F.err = c->d_counters + 6;
        uint32_t* he = hf + 1;
        if (!check(c, lyg_launch_fill_u32(c->d_counters + 2, 1, 0, s), "reset") ||
            !check(c, lyg_launch_fill_u32(c->d_counters + 6, 1, 0, s), "reset") ||
            !check(c, lyg_launch_stroke(&F, 1, persistent_grid(c, *hf), s), "stroke emit (again)") ||
            !check(c, hipMemcpyAsync(he, c->d_counters + 6, 4, hipMemcpyDeviceToHost, s), "copy") ||
            !check(c, hipStreamSynchronize(s), "sync"))
            return LYG_ERR_DEVICE;
        if (*he) {
            c->last_error = "emit pass wrote outside its path range (staged/final mismatch)";
            return LYG_ERR_INTERNAL;
        }
    }
    c->last_stroke_items = n_items;
    c->last_stroke_again = *hf;
    if (trace_on()) {
        std::vector<lyg::StrokeHeads> hh(n);
        (void)hipMemcpy(hh.data(), L.heads, sizeof(lyg::StrokeHeads) * n, hipMemcpyDeviceToHost);
        uint32_t split = 0, mism = 0, over = 0;
        for (const auto& x : hh) {
            split += x.n_items > 1;
            mism += x.redo == 1 || x.redo == 2;
            over += x.redo == 3;
        }
        fprintf(stderr, "[lyg] stroke items: %u paths (%u split), %llu items, staging %llu v / "
                        "%llu i, out %llu v / %llu i, %u emitted again (%u state mismatches, "
                        "%u rooms outgrown)\n",
                n, split, (unsigned long long)n_items, (unsigned long long)sv,
                (unsigned long long)si, (unsigned long long)tv, (unsigned long long)ti, *hf, mism,
                over);
        fflush(stderr);
    }
    return LYG_OK;
}

}  // namespace

int32_t lygi::stroke_batch(lyg_ctx* c, const lygi::BatchSpec& b, const lygi::OutAlloc& alloc,
                           uint64_t* tv_out, uint64_t* ti_out) {
    const uint32_t n = b.in.n_paths;
    if (!ensure(c, c->counts_v, 4ull * n) || !ensure(c, c->counts_i, 4ull * n))
        return LYG_ERR_DEVICE;
    lyg::StrokeLaunch L{};
    L.nsub_thr = as<float>(c->nsub_thr);
    L.in = b.in;
    L.opts = reinterpret_cast<const lyg_stroke_options*>(b.opts);
    L.opts_stride = b.opts_stride;
    L.path_nv = as<uint32_t>(c->counts_v);
    L.path_ni = as<uint32_t>(c->counts_i);
    L.vertex_offset = b.vertex_offset;
    L.local_ids = b.local_ids ? 1u : 0u;
    if (c->stroke_staged && n > 0) {
        lyg::StrokeLaunch Ls = L;
        const int32_t r = c->stroke_split_windows > 0
                              ? stroke_batch_items(c, b, alloc, Ls, tv_out, ti_out)
                              : stroke_batch_staged(c, b, alloc, Ls, tv_out, ti_out);
        if (r != 1) return r;
    }
    uint64_t tv = 0, ti = 0;
    int32_t st = stroke_count_and_scan(c, L, b.vtx_off, b.idx_off, &tv, &ti);
    *tv_out = tv;
    *ti_out = ti;
    if (st != LYG_OK) return st;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (!b.local_ids && tv > 0 && (uint64_t)b.vertex_offset + tv - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    st = alloc(tv, ti, &ov, &oi);
    if (st != LYG_OK) return st;
    L.vtx_off = b.vtx_off;
    L.idx_off = b.idx_off;
    L.out_vertices = ov;
    L.out_indices = oi;
    L.vertex_capacity = tv;
    L.index_capacity = ti;
    st = stroke_emit(c, L);
    if (st != LYG_OK) return st;
    if (!check(c, hipStreamSynchronize(c->stream), "sync")) return LYG_ERR_DEVICE;
    if (emit_failed(c)) return LYG_ERR_INTERNAL;
    return LYG_OK;
}

namespace {

bool valid_stroke_opts(const lyg_stroke_options& o) {
    return o.start_cap <= 2 && o.end_cap <= 2 && o.line_join <= 3;
}

}  // namespace

extern "C" {

lyg_ctx* lyg_ctx_create(int device) { return lygi::ctx_create(device, true); }

}  // extern "C"

lyg_ctx* lygi::ctx_create(int device, bool side_streams) {
    if (trace_on()) {
        fprintf(stderr, "[lyg] ctx_create(%d)\n", device);
        fflush(stderr);
    }
    lyg_ctx* c = new lyg_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete c;
        return nullptr;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return nullptr;
    }
    // The side streams of the fill's concurrent launches are made here, right after the
    // context's own stream, so that they get hardware queues of their own (a process has
    // four): streams made later (a HIP graph's replay makes some) would share queues with
    // them and serialise the concurrent launches.
    for (int k = 0; side_streams && k < 3; ++k)
        if (hipStreamCreateWithFlags(&c->class_stream[k], hipStreamNonBlocking) != hipSuccess) {
            lyg_ctx_destroy(c);
            return nullptr;
        }
    if (hipMalloc(&c->d_counters, 64) != hipSuccess || hipMalloc(&c->d_totals, 64) != hipSuccess ||
        hipHostMalloc(&c->h_pinned, 256, hipHostMallocDefault) != hipSuccess ||
        lyg_noalloc_scratch_init(device) != hipSuccess) {
        lyg_ctx_destroy(c);
        return nullptr;
    }
    // round-join subdivision thresholds, once per context
    if (!ensure(c, c->nsub_thr, 32 * sizeof(float)) ||
        lyg_launch_nsub_thresholds(as<float>(c->nsub_thr), c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
        lyg_ctx_destroy(c);
        return nullptr;
    }
    return c;
}

extern "C" {

void lyg_ctx_destroy(lyg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->in_verbs, &c->in_points, &c->in_offs, &c->in_opts, &c->counts_v,
                      &c->counts_i, &c->offs_v,    &c->offs_i,  &c->out_v,   &c->out_i,
                      &c->misc,     &c->arcs,      &c->flat_kinds, &c->flat_ctrl,
                      &c->flat_runs, &c->flat_poly, &c->in_attrs, &c->out_attrs};
    for (DevBuf* b : bufs) free_buf(*b);
    free_buf(c->f_bound);
    free_buf(c->f_status);
    free_buf(c->f_round);
    free_buf(c->f_hint);
    free_buf(c->f_work);
    free_buf(c->f_inter);
    free_buf(c->nsub_thr);
    free_buf(c->scan_tiles);
    for (DevBuf* b : {&c->st_cap_v, &c->st_cap_i, &c->st_off_v, &c->st_off_i, &c->st_v, &c->st_i,
                      &c->st_fallback, &c->st_items, &c->st_heads, &c->st_spec, &c->st_end,
                      &c->st_parts, &c->st_parts_off})
        free_buf(*b);
    free_buf(c->f_boxes);
    free_buf(c->f_counters);
    lyg_band_free(c);
    for (auto* v : {&c->f_stage_v, &c->f_stage_t, &c->f_stage_off, &c->f_list})
        for (DevBuf& b : *v) free_buf(b);
    if (c->d_counters) (void)hipFree(c->d_counters);
    if (c->d_totals) (void)hipFree(c->d_totals);
    if (c->h_pinned) (void)hipHostFree(c->h_pinned);
    for (int i = 0; i < 4; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    for (hipStream_t& cs : c->class_stream)
        if (cs) (void)hipStreamDestroy(cs);
    for (hipEvent_t& ce : c->class_ev)
        if (ce) (void)hipEventDestroy(ce);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* lyg_status_string(int32_t s) {
    switch (s) {
        case LYG_OK: return "ok";
        case LYG_ERR_TOO_MANY_VERTICES: return "GeometryBuilderError::TooManyVertices";
        case LYG_ERR_INVALID_VERTEX: return "GeometryBuilderError::InvalidVertex";
        case LYG_ERR_POSITION_IS_NAN: return "UnsupportedParamater::PositionIsNaN";
        case LYG_ERR_TOLERANCE_IS_NAN: return "UnsupportedParamater::ToleranceIsNaN";
        case LYG_ERR_INTERNAL: return "InternalError";
        case LYG_ERR_CAPACITY: return "output capacity too small";
        case LYG_ERR_INVALID_ARGUMENT: return "invalid argument";
        case LYG_ERR_DEVICE: return "HIP runtime error";
        case LYG_ERR_UNSUPPORTED: return "unsupported option";
        case LYG_ERR_NO_SPACE: return "no space in the device pool";
        default: return "unknown";
    }
}

const char* lyg_ctx_last_error(const lyg_ctx* c) { return c ? c->last_error.c_str() : ""; }

int32_t lyg_ctx_enable_timing(lyg_ctx* c, int enable) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    if (enable && !c->ev[0]) {
        for (int i = 0; i < 4; ++i)
            if (!check(c, hipEventCreate(&c->ev[i]), "hipEventCreate")) return LYG_ERR_DEVICE;
    }
    c->timing = enable != 0;
    return LYG_OK;
}

int32_t lyg_ctx_kernel_times(lyg_ctx* c, float* count_ms, float* emit_ms) {
    if (!c || !c->timing) return LYG_ERR_INVALID_ARGUMENT;
    if (!check(c, hipEventSynchronize(c->ev[3]), "hipEventSynchronize")) return LYG_ERR_DEVICE;
    float a = 0, b = 0;
    if (!check(c, hipEventElapsedTime(&a, c->ev[0], c->ev[1]), "elapsed") ||
        !check(c, hipEventElapsedTime(&b, c->ev[2], c->ev[3]), "elapsed"))
        return LYG_ERR_DEVICE;
    if (count_ms) *count_ms = a;
    if (emit_ms) *emit_ms = b;
    return LYG_OK;
}

// ---- device memory --------------------------------------------------------------------
int32_t lyg_device_alloc(lyg_ctx* c, uint64_t bytes, void** ptr) {
    if (!c || !ptr) return LYG_ERR_INVALID_ARGUMENT;
    *ptr = nullptr;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    if (!check(c, hipMalloc(ptr, bytes ? bytes : 16), "hipMalloc")) return LYG_ERR_NO_SPACE;
    return LYG_OK;
}

int32_t lyg_device_free(lyg_ctx* c, void* ptr) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    if (!ptr) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice") || !check(c, hipFree(ptr), "hipFree"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_copy_to_device(lyg_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return LYG_ERR_INVALID_ARGUMENT;
    if (!bytes) return LYG_OK;
    if (!check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "H2D") ||
        !check(c, hipStreamSynchronize(c->stream), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_copy_to_host(lyg_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return LYG_ERR_INVALID_ARGUMENT;
    if (!bytes) return LYG_OK;
    if (!check(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "D2H") ||
        !check(c, hipStreamSynchronize(c->stream), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_device_copy(lyg_ctx* c, void* dst, const void* src, uint64_t bytes) {
    if (!c || (bytes && (!dst || !src))) return LYG_ERR_INVALID_ARGUMENT;
    if (!bytes) return LYG_OK;
    if (((uintptr_t)dst | (uintptr_t)src) & 15u) return LYG_ERR_INVALID_ARGUMENT;
    const int grid = c->cu_count * 8;
    if (!check(c, lyg_launch_copy(dst, src, bytes, grid, c->stream), "copy"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_ctx_synchronize(lyg_ctx* c) {
    if (!c) return LYG_ERR_INVALID_ARGUMENT;
    return check(c, hipStreamSynchronize(c->stream), "sync") ? LYG_OK : LYG_ERR_DEVICE;
}

// ---- plain per-path stroke -------------------------------------------------------------
int32_t lyg_stroke_path(lyg_ctx* c, const lyg_path_view* path, const lyg_stroke_options* opts,
                        lyg_output* out) {
    if (!c || !path || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (!valid_stroke_opts(*opts)) return LYG_ERR_INVALID_ARGUMENT;
    if ((path->n_verbs && !path->verbs) || (path->n_points && !path->points))
        return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (path->n_verbs == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    if (!ensure(c, c->in_verbs, path->n_verbs) || !ensure(c, c->in_points, 8ull * path->n_points) ||
        !ensure(c, c->in_offs, 16) || !ensure(c, c->in_opts, sizeof(lyg_stroke_options)) ||
        !ensure(c, c->counts_v, 8) || !ensure(c, c->counts_i, 8) || !ensure(c, c->offs_v, 8) ||
        !ensure(c, c->offs_i, 8))
        return LYG_ERR_DEVICE;
    uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 64);
    h[0] = 0;
    h[1] = path->n_verbs;
    h[2] = 0;
    h[3] = path->n_points;
    hipStream_t s = c->stream;
    if (!check(c, hipMemcpyAsync(c->in_verbs.p, path->verbs, path->n_verbs, hipMemcpyHostToDevice, s), "H2D") ||
        (path->n_points &&
         !check(c, hipMemcpyAsync(c->in_points.p, path->points, 8ull * path->n_points, hipMemcpyHostToDevice, s), "H2D")) ||
        !check(c, hipMemcpyAsync(c->in_offs.p, h, 16, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipMemcpyAsync(c->in_opts.p, opts, sizeof(*opts), hipMemcpyHostToDevice, s), "H2D"))
        return LYG_ERR_DEVICE;
    lyg::StrokeLaunch L{};
    L.nsub_thr = as<float>(c->nsub_thr);
    L.in.verbs = as<uint8_t>(c->in_verbs);
    L.in.points = as<float>(c->in_points);
    L.in.verb_off = as<uint32_t>(c->in_offs);
    L.in.point_off = as<uint32_t>(c->in_offs) + 2;
    L.in.n_paths = 1;
    L.in.first_path = 0;
    L.opts = as<lyg_stroke_options>(c->in_opts);
    L.opts_stride = 0;
    L.path_nv = as<uint32_t>(c->counts_v);
    L.path_ni = as<uint32_t>(c->counts_i);
    L.vertex_offset = out->vertex_offset;
    uint64_t tv = 0, ti = 0;
    int32_t st = stroke_count_and_scan(c, L, as<uint32_t>(c->offs_v), as<uint32_t>(c->offs_i), &tv, &ti);
    if (st != LYG_OK) return st;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350).
    if (tv > 0 && (uint64_t)out->vertex_offset + tv - 1 > 0xfffffffeull) return LYG_ERR_TOO_MANY_VERTICES;
    if (tv > out->vertex_capacity || ti > out->index_capacity) {
        out->n_vertices = (uint32_t)tv;
        out->n_indices = (uint32_t)ti;
        return LYG_ERR_CAPACITY;
    }
    if (tv == 0 && ti == 0) return LYG_OK;
    if (!ensure(c, c->out_v, 8 * tv) || !ensure(c, c->out_i, 4 * ti)) return LYG_ERR_DEVICE;
    L.vtx_off = as<uint32_t>(c->offs_v);
    L.idx_off = as<uint32_t>(c->offs_i);
    L.out_vertices = as<float>(c->out_v);
    L.out_indices = as<uint32_t>(c->out_i);
    L.vertex_capacity = tv;
    L.index_capacity = ti;
    st = stroke_emit(c, L);
    if (st != LYG_OK) return st;
    if (!check(c, hipMemcpyAsync(out->vertices, c->out_v.p, 8 * tv, hipMemcpyDeviceToHost, s), "D2H") ||
        (ti && !check(c, hipMemcpyAsync(out->indices, c->out_i.p, 4 * ti, hipMemcpyDeviceToHost, s), "D2H")) ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    if (emit_failed(c)) return LYG_ERR_INTERNAL;
    out->n_vertices = (uint32_t)tv;
    out->n_indices = (uint32_t)ti;
    return LYG_OK;
}

// ---- batched stroke (device-resident) -------------------------------------------------
int32_t lyg_stroke_batch_count(lyg_ctx* c, const lyg_batch_in* in, const lyg_stroke_options* opts,
                               uint32_t n_opts, uint32_t* path_nv, uint32_t* path_ni,
                               uint64_t* total_nv, uint64_t* total_ni) {
    if (!c || !in || !opts || !path_nv || !path_ni) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    if (in->n_paths == 0) {
        if (total_nv) *total_nv = 0;
        if (total_ni) *total_ni = 0;
        return LYG_OK;
    }
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lyg::StrokeLaunch L{};
    L.nsub_thr = as<float>(c->nsub_thr);
    L.in.verbs = in->verbs;
    L.in.points = in->points;
    L.in.verb_off = in->verb_offsets;
    L.in.point_off = in->point_offsets;
    L.in.n_paths = in->n_paths;
    L.opts = opts;
    L.opts_stride = n_opts == 1 ? 0 : 1;
    L.path_nv = path_nv;
    L.path_ni = path_ni;
    if (!ensure(c, c->offs_v, 4ull * (in->n_paths + 1)) || !ensure(c, c->offs_i, 4ull * (in->n_paths + 1)))
        return LYG_ERR_DEVICE;
    uint64_t tv, ti;
    int32_t st = stroke_count_and_scan(c, L, as<uint32_t>(c->offs_v), as<uint32_t>(c->offs_i), &tv, &ti);
    if (total_nv) *total_nv = tv;
    if (total_ni) *total_ni = ti;
    return st;
}

int32_t lyg_stroke_batch(lyg_ctx* c, const lyg_batch_in* in, const lyg_stroke_options* opts,
                         uint32_t n_opts, lyg_batch_out* out) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (in->n_paths == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lygi::BatchSpec b{};
    b.in.verbs = in->verbs;
    b.in.points = in->points;
    b.in.verb_off = in->verb_offsets;
    b.in.point_off = in->point_offsets;
    b.in.n_paths = in->n_paths;
    b.opts = opts;
    b.opts_stride = n_opts == 1 ? 0 : 1;
    b.vertex_offset = out->vertex_offset;
    b.local_ids = false;
    b.vtx_off = out->path_vertex_offsets;
    b.idx_off = out->path_index_offsets;
    uint64_t tv = 0, ti = 0;
    int32_t st = lygi::stroke_batch(
        c, b,
        [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
            if (nv > out->vertex_capacity || ni > out->index_capacity) return LYG_ERR_CAPACITY;
            *v = out->vertices;
            *i = out->indices;
            return LYG_OK;
        },
        &tv, &ti);
    out->n_vertices = tv;
    out->n_indices = ti;
    if (st != LYG_OK) return st;
    if (out->path_status &&
        !check(c, lyg_launch_fill_u32(reinterpret_cast<uint32_t*>(out->path_status), in->n_paths, 0,
                                      c->stream), "status"))
        return LYG_ERR_DEVICE;
    if (!check(c, hipStreamSynchronize(c->stream), "sync")) return LYG_ERR_DEVICE;
    return LYG_OK;
}

// ---- flattening ---------------------------------------------------------------------------
int32_t lyg_flatten_segments(lyg_ctx* c, const int32_t* kinds, const float* ctrl, uint32_t n,
                             float tolerance, float* out_points, uint32_t point_capacity,
                             uint32_t* offsets, uint32_t* n_points) {
    if (!c || (n && (!kinds || !ctrl || !offsets))) return LYG_ERR_INVALID_ARGUMENT;
    if (!(tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (n_points) *n_points = 0;
    if (n == 0) {
        offsets[0] = 0;
        return LYG_OK;
    }
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    if (!ensure(c, c->flat_kinds, 4ull * n) || !ensure(c, c->flat_ctrl, 32ull * n) ||
        !ensure(c, c->counts_v, 4ull * n) || !ensure(c, c->offs_v, 4ull * (n + 1)))
        return LYG_ERR_DEVICE;
    if (!check(c, hipMemcpyAsync(c->flat_kinds.p, kinds, 4ull * n, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipMemcpyAsync(c->flat_ctrl.p, ctrl, 32ull * n, hipMemcpyHostToDevice, s), "H2D"))
        return LYG_ERR_DEVICE;
    lyg::FlattenLaunch L{};
    L.kinds = as<int32_t>(c->flat_kinds);
    L.ctrl = as<float>(c->flat_ctrl);
    L.n = n;
    L.tolerance = tolerance;
    L.counts = as<uint32_t>(c->counts_v);
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, lyg_launch_flatten_count(&L, s), "flatten count") ||
        !check(c, ctx_scan(c, L.counts, n, as<uint32_t>(c->offs_v), c->d_totals, c->d_counters + 4, s), "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    uint64_t total = h[0];
    if (n_points) *n_points = (uint32_t)total;
    if (total > point_capacity || !out_points) return LYG_ERR_CAPACITY;
    if (!ensure(c, c->out_v, 8 * total) ||
        !ensure(c, c->flat_runs, lyg_flatten_runs_bytes((uint32_t)total)) ||
        !ensure(c, c->flat_poly, lyg_flatten_poly_bytes(n)))
        return LYG_ERR_DEVICE;
    L.offsets = as<uint32_t>(c->offs_v);
    L.out_points = as<float>(c->out_v);
    L.runs = as<uint32_t>(c->flat_runs);
    L.polys = as<float>(c->flat_poly);
    if (!check(c, lyg_launch_flatten_emit(&L, (uint32_t)total, s), "flatten emit") ||
        !check(c, hipMemcpyAsync(out_points, c->out_v.p, 8 * total, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(offsets, c->offs_v.p, 4ull * (n + 1), hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_flatten_segments_device(lyg_ctx* c, const int32_t* kinds, const float* ctrl,
                                    uint32_t n, float tolerance, float* out_points,
                                    uint32_t point_capacity, uint32_t* offsets,
                                    uint32_t* n_points) {
    if (!c || (n && (!kinds || !ctrl || !offsets))) return LYG_ERR_INVALID_ARGUMENT;
    if (!(tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (n_points) *n_points = 0;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    if (n == 0) {
        return check(c, lyg_launch_fill_u32(offsets, 1, 0, s), "zero") &&
                       check(c, hipStreamSynchronize(s), "sync")
                   ? LYG_OK
                   : LYG_ERR_DEVICE;
    }
    if (!ensure(c, c->counts_v, 4ull * n)) return LYG_ERR_DEVICE;
    lyg::FlattenLaunch L{};
    L.kinds = kinds;
    L.ctrl = ctrl;
    L.n = n;
    L.tolerance = tolerance;
    L.counts = as<uint32_t>(c->counts_v);
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, lyg_launch_flatten_count(&L, s), "flatten count") ||
        !check(c, ctx_scan(c, L.counts, n, offsets, c->d_totals, c->d_counters + 4, s), "scan"))
        return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 1, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t total = h[0];
    if ((uint32_t)h[1] != 0 || total > 0xffffffffull) return LYG_ERR_TOO_MANY_VERTICES;
    if (n_points) *n_points = (uint32_t)total;
    if (total > point_capacity || !out_points) return LYG_ERR_CAPACITY;
    if (!ensure(c, c->flat_runs, lyg_flatten_runs_bytes((uint32_t)total)) ||
        !ensure(c, c->flat_poly, lyg_flatten_poly_bytes(n)))
        return LYG_ERR_DEVICE;
    L.offsets = offsets;
    L.out_points = out_points;
    L.runs = as<uint32_t>(c->flat_runs);
    L.polys = as<float>(c->flat_poly);
    if (c->timing) (void)hipEventRecord(c->ev[2], s);
    if (!check(c, lyg_launch_flatten_emit(&L, (uint32_t)total, s), "flatten emit")) return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[3], s);
    if (!check(c, hipStreamSynchronize(s), "sync")) return LYG_ERR_DEVICE;
    return LYG_OK;
}

// lyg_flatten_segments_device without any allocation: the counts, the scan's tile sums,
// the totals and the overflow flag live in the library's fixed per-device scratch array
// (set up by lyg_ctx_create); the host reads the total through the context's pinned
// staging.  No hipMalloc / hipFree / allocating library call on this path.
int32_t lyg_flatten_segments_device_noalloc(lyg_ctx* c, const int32_t* kinds, const float* ctrl,
                                            uint32_t n, float tolerance, float* out_points,
                                            uint32_t point_capacity, uint32_t* offsets,
                                            uint32_t* n_points) {
    if (!c || (n && (!kinds || !ctrl || !offsets))) return LYG_ERR_INVALID_ARGUMENT;
    if (!(tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (n_points) *n_points = 0;
    if (n > lyg_noalloc_max_segments()) {
        c->last_error = "lyg_flatten_segments_device_noalloc: more segments than its fixed "
                        "scratch holds (lyg_noalloc_max_segments)";
        return LYG_ERR_CAPACITY;
    }
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    uint32_t *counters, *counts;
    uint64_t *totals, *tiles;
    lyg_noalloc_scratch(c->device, &counters, &totals, &counts, &tiles);
    if (!counters) {
        c->last_error = "allocation-free scratch not set up on this device";
        return LYG_ERR_DEVICE;
    }
    hipStream_t s = c->stream;
    if (n == 0) {
        return check(c, lyg_launch_fill_u32(offsets, 1, 0, s), "zero") &&
                       check(c, hipStreamSynchronize(s), "sync")
                   ? LYG_OK
                   : LYG_ERR_DEVICE;
    }
    lyg::FlattenLaunch L{};
    L.kinds = kinds;
    L.ctrl = ctrl;
    L.n = n;
    L.tolerance = tolerance;
    L.counts = counts;
    // the scratch keeps the previous call's contents: the overflow flag must start at zero
    if (!check(c, lyg_launch_fill_u32(counters, 16, 0, s), "reset") ||
        !check(c, lyg_launch_flatten_count(&L, s), "flatten count") ||
        !check(c, lyg_launch_scan_tiles(counts, n, offsets, totals, counters, tiles, s), "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, totals, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 1, counters, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t total = h[0];
    if ((uint32_t)h[1] != 0 || total > 0xffffffffull) return LYG_ERR_TOO_MANY_VERTICES;
    if (n_points) *n_points = (uint32_t)total;
    if (total > point_capacity || !out_points) return LYG_ERR_CAPACITY;
    L.offsets = offsets;
    L.out_points = out_points;
    if (!check(c, lyg_launch_flatten_emit(&L, (uint32_t)total, s), "flatten emit") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

uint32_t lyg_flatten_noalloc_max_segments(void) { return lyg_noalloc_max_segments(); }

static int32_t arcs_common(lyg_ctx* c, const float* arcs, uint32_t n, float tol, bool quads,
                           float* out, uint32_t cap, uint32_t* offsets, uint32_t* n_out) {
    if (!c || (n && (!arcs || !offsets))) return LYG_ERR_INVALID_ARGUMENT;
    if (n_out) *n_out = 0;
    if (n == 0) {
        offsets[0] = 0;
        return LYG_OK;
    }
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    if (!ensure(c, c->arcs, 28ull * n) || !ensure(c, c->counts_v, 4ull * n) ||
        !ensure(c, c->offs_v, 4ull * (n + 1)))
        return LYG_ERR_DEVICE;
    if (!check(c, hipMemcpyAsync(c->arcs.p, arcs, 28ull * n, hipMemcpyHostToDevice, s), "H2D"))
        return LYG_ERR_DEVICE;
    hipError_t e = quads ? lyg_launch_arc_to_quads(as<float>(c->arcs), n, as<uint32_t>(c->counts_v), nullptr, nullptr, s)
                         : lyg_launch_arc_flatten(as<float>(c->arcs), n, tol, as<uint32_t>(c->counts_v), nullptr, nullptr, s);
    if (!check(c, e, "arc count") ||
        !check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, ctx_scan(c, as<uint32_t>(c->counts_v), n, as<uint32_t>(c->offs_v), c->d_totals, c->d_counters + 4, s), "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    uint64_t total = h[0];
    if (n_out) *n_out = (uint32_t)total;
    if (total > cap || !out) return LYG_ERR_CAPACITY;
    size_t per = quads ? 24 : 8;
    if (!ensure(c, c->out_v, per * total)) return LYG_ERR_DEVICE;
    e = quads ? lyg_launch_arc_to_quads(as<float>(c->arcs), n, nullptr, as<uint32_t>(c->offs_v), as<float>(c->out_v), s)
              : lyg_launch_arc_flatten(as<float>(c->arcs), n, tol, nullptr, as<uint32_t>(c->offs_v), as<float>(c->out_v), s);
    if (!check(c, e, "arc emit") ||
        !check(c, hipMemcpyAsync(out, c->out_v.p, per * total, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(offsets, c->offs_v.p, 4ull * (n + 1), hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_flatten_arcs(lyg_ctx* c, const float* arcs, uint32_t n, float tolerance,
                         float* out_points, uint32_t point_capacity, uint32_t* offsets,
                         uint32_t* n_points) {
    if (!(tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
    return arcs_common(c, arcs, n, tolerance, false, out_points, point_capacity, offsets, n_points);
}

int32_t lyg_arc_to_quadratics(lyg_ctx* c, const float* arcs, uint32_t n, float* out_quads,
                              uint32_t quad_capacity, uint32_t* offsets, uint32_t* n_quads) {
    return arcs_common(c, arcs, n, 0.0f, true, out_quads, quad_capacity, offsets, n_quads);
}

}  // extern "C"
