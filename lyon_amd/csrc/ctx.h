// Internal to liblyg.so: the context object and the helpers every host file shares.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/lyg.h"

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

struct BandBufs;  // the banded fill's buffers (band_api.cpp)

struct lyg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;
    std::string last_error;
    // grow-only device scratch
    DevBuf in_verbs, in_points, in_offs, in_opts, counts_v, counts_i, offs_v, offs_i, out_v,
        out_i, misc, arcs, flat_kinds, flat_ctrl, flat_runs, flat_poly, in_attrs, out_attrs;
    uint32_t* d_counters = nullptr;  // [0..3] work-queue heads, [4] overflow flag
    uint64_t* d_totals = nullptr;    // [0] vertices, [1] indices
    void* h_pinned = nullptr;        // small pinned staging area for totals
    // optional kernel timing: events around the count and emit kernels on `stream`
    bool timing = false;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // stroke launch shape (set by the autotuner): resident blocks per CU of the grid
    uint32_t stroke_blocks_per_cu = 8;
    // fill launch shape (set by the autotuner): waves per block, LDS bytes per wave
    uint32_t fill_waves_per_block = 2;
    uint32_t fill_hot_bytes = 16 * 1024;
    uint32_t fill_lds_resident = 1;  // first round in the LDS-resident form where it fits
    // fill: per-path arrays, per-wave working memory and per-round staging
    DevBuf f_bound, f_status, f_round, f_hint, f_work, f_inter, f_boxes, f_counters;
    std::vector<DevBuf> f_stage_v, f_stage_t, f_stage_off, f_list;
    // stroke: staged single pass (emit into per-path staging room, then compact)
    bool stroke_staged = true;
    float stroke_room_scale = 1.0f;  // multiplies every path's staging room
    DevBuf nsub_thr;                 // round-join subdivision thresholds (31 floats)
    DevBuf scan_tiles;               // tile sums of scans over more than 32768 counts
    DevBuf st_cap_v, st_cap_i, st_off_v, st_off_i, st_v, st_i, st_fallback;
    // stroke: long paths split into parts of this many 64-step windows (0: whole paths)
    uint32_t stroke_split_windows = 64;
    uint64_t last_stroke_items = 0;       // items of the last split stroke batch
    uint32_t last_stroke_again = 0;       // its paths walked again whole
    DevBuf st_items, st_heads, st_spec, st_end, st_parts, st_parts_off;
    // fill: one stream per size class of a round, and the events that join them
    hipStream_t class_stream[8] = {};
    hipEvent_t class_ev[25] = {};  // one per fill launch (size class x form) + 1
    // banded fill: band size (edge starts per band) and the most bands of one path
    BandBufs* band = nullptr;
    uint32_t band_edges_per_band = 256;
    uint32_t band_max_bands = 16384;
};

void lyg_band_free(lyg_ctx* c);

inline void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// LYG_TRACE=1 prints every host step to stderr (flushed), for diagnosing a run that
// goes quiet.
inline bool trace_on() {
    static int on = -1;
    if (on < 0) {
        const char* v = getenv("LYG_TRACE");
        on = (v && *v && *v != '0') ? 1 : 0;
    }
    return on == 1;
}

inline bool check(lyg_ctx* c, hipError_t e, const char* what) {
    if (trace_on()) {
        fprintf(stderr, "[lyg] %s -> %d\n", what, (int)e);
        fflush(stderr);
    }
    if (e == hipSuccess) return true;
    c->last_error = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

inline bool ensure(lyg_ctx* c, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return true;
    const size_t old_cap = b.cap;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    // Buffers below 256 MiB at least double when they grow: a growth frees the old buffer,
    // and hipFree waits for the whole device, stalling every stream of the process (the
    // serving workers' batches grow their buffers in many small steps otherwise).
    size_t cap = std::max(bytes, (size_t)4096);
    if (bytes < ((size_t)256 << 20)) cap = std::max(cap, 2 * old_cap);
    cap = (cap + 4095) & ~(size_t)4095;
    if (!check(c, hipMalloc(&b.p, cap), "hipMalloc")) return false;
    b.cap = cap;
    return true;
}

template <class T>
T* as(DevBuf& b) {
    return reinterpret_cast<T*>(b.p);
}

inline int persistent_grid(lyg_ctx* c, uint32_t n_paths) {
    // 4 waves per block, one path per wave at a time; stroke_blocks_per_cu blocks per CU
    // (the autotuner's choice; 8 keeps 32 waves per CU).
    uint32_t blocks_needed = (n_paths + 3) / 4;
    uint32_t cap = (uint32_t)c->cu_count * c->stroke_blocks_per_cu;
    return (int)std::max(1u, std::min(blocks_needed, cap));
}

extern "C" {
hipError_t lyg_launch_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total,
                           uint32_t* overflow, hipStream_t s);
size_t lyg_scan_tiles_bytes(uint32_t n);
hipError_t lyg_launch_scan_tiles(const uint32_t* counts, uint32_t n, uint32_t* offsets,
                                 uint64_t* total, uint32_t* overflow, uint64_t* tiles,
                                 hipStream_t s);
}

// Exclusive scan of n counts into offsets (n + 1) and *total: one block for small n, the
// whole GPU (with the tile buffer `tiles`) beyond 32768 counts.  A captured graph passes a
// buffer of its own: a later call can reallocate the context's.
inline hipError_t scan_with(lyg_ctx* c, DevBuf& tiles, const uint32_t* counts, uint32_t n,
                            uint32_t* offsets, uint64_t* total, uint32_t* overflow,
                            hipStream_t s) {
    if (n > 32768) {
        if (!ensure(c, tiles, lyg_scan_tiles_bytes(n))) return hipErrorOutOfMemory;
        return lyg_launch_scan_tiles(counts, n, offsets, total, overflow,
                                     static_cast<uint64_t*>(tiles.p), s);
    }
    return lyg_launch_scan(counts, n, offsets, total, overflow, s);
}
inline hipError_t ctx_scan(lyg_ctx* c, const uint32_t* counts, uint32_t n, uint32_t* offsets,
                           uint64_t* total, uint32_t* overflow, hipStream_t s) {
    return scan_with(c, c->scan_tiles, counts, n, offsets, total, overflow, s);
}
