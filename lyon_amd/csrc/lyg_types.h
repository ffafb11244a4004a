// Plain-old-data layouts shared by the host runtime and the HIP kernels.
#pragma once
#include <stdint.h>

#include "../../include/lyg.h"

namespace lyg {

// Device-side description of one batch of paths, in lyon's flat storage layout
// (path.rs:75-79): verbs[] and points[] of every path back to back, with n+1 offsets.
struct PathBatchDev {
    const uint8_t* verbs;
    const float* points;  // x, y pairs
    const uint32_t* verb_off;
    const uint32_t* point_off;
    uint32_t n_paths;
    uint32_t first_path;  // index of the first path of this launch within the batch
};

// ---- long strokes split into parts walked at once --------------------------------------
// A part of a path's walk: windows [ke, we) .. (k1, w1) of the path's steps, after a
// prologue from (k0, w0) that rebuilds the carry.  Parts of one path are walked by
// different waves; a check afterwards confirms that every part's prologue rebuilt the
// state the previous part ended in, else the path is walked again whole.
struct StrokeItem {
    uint32_t path;
    uint32_t flags;                  // ITEM_* below
    uint32_t k0, w0, ke, we, k1, w1; // verb chunk / step of the prologue, start and end
    uint32_t pt0;                    // points consumed before chunk k0
    uint32_t v_off, v_room;          // staging room (vertices), item-local ids
    uint32_t i_off, i_room;          // staging room (index slots)
    uint32_t nv, ni;                 // exact counts (vertices, index slots)
    uint32_t vbase, ibase;           // the item's first vertex / index slot in its path
};
enum : uint32_t { ITEM_FIRST = 1u, ITEM_SPLIT = 2u };
// Ids with this bit pattern (0b110 in the top three bits) refer to the first vertices of
// the path, not of the item: the last part of a split path writes them at the sub-path's
// end.  The range is apart from the item-local ids below 0x80000000 and from the stand-in
// ids the walk counts new vertices with (0x80000000 + lane << 24 + k) before their real
// ids are known: the walk compares ids for equality.
static constexpr uint32_t ID_HEAD_MARK = 0xc0000000u;
static constexpr uint32_t ID_HEAD_MASK = 0xe0000000u;
// Item-local ids start here, so the ids a part refers to in the part before it (a few
// vertices back) stay below it and far from INVALID_ID: the walk compares ids for
// equality, and an id that wrapped onto 0xffffffff would read as "no vertex".
static constexpr uint32_t ITEM_ID_BIAS = 0x40000000u;

struct F2 {
    float x, y;
};

// Per path: its first points (the first window walked alone), its items, the check.
struct StrokeHeads {
    F2 f0, f1, f2;
    uint32_t ff1, f1_base;
    uint32_t n_items, extra_first;  // item 0 is item `path`; items 1.. at extra_first..
    uint32_t redo;                  // 1: the parts disagreed or outgrew their room
};

// The carry state that decides the rest of a walk (see carry_store in stroke.hip).
struct CarryRec {
    F2 pos[3];
    F2 next[3][2];
    F2 last_kept;
    uint32_t flags[3], isp[3], code[3], base[3];
    uint32_t sp, prev_end, mnec, pad;
};

struct StrokeLaunch {
    PathBatchDev in;
    const lyg_stroke_options* opts;  // per path (opts_stride 1) or shared (stride 0)
    uint32_t opts_stride;
    // count pass outputs
    uint32_t* path_nv;
    uint32_t* path_ni;
    // emit pass inputs/outputs
    const uint32_t* vtx_off;  // exclusive offsets (n + 1)
    const uint32_t* idx_off;
    float* out_vertices;
    uint32_t* out_indices;
    uint32_t vertex_offset;  // BuffersBuilder::with_vertex_offset
    uint32_t local_ids;      // ids count from vertex_offset in every path (not the batch)
    uint32_t* work_counter;  // device work queue head (persistent kernels)
    uint32_t* err;           // emit pass: bit 0 a write outside its path's range, bit 1 a
                             // path beyond the output capacity (the path is skipped)
    uint64_t vertex_capacity;  // emit pass: output array sizes
    uint64_t index_capacity;
    const uint32_t* path_list;  // optional: the launch walks path_list[0 .. n_paths)
    uint32_t store_counts;      // emit pass also stores per-path counts (staged emit)
    const float* nsub_thr;      // round-join subdivision thresholds (31 floats) or null
    // split walks (null items: one wave walks each whole path)
    StrokeItem* items;          // n_paths + extra items; the work queue runs over n_items
    uint32_t n_items;
    // graph replay: the item count is n_paths + min(*n_extra_dev, cap_extra), and a
    // path-list walk takes its length from *n_list_dev (null: the host values above)
    const unsigned long long* n_extra_dev;
    uint32_t cap_extra;
    const uint32_t* n_list_dev;
    StrokeHeads* heads;         // per path
    CarryRec* item_spec;        // per item: the carry its prologue rebuilt
    CarryRec* item_end;         // per item: the carry at its end
};

// Staged single-pass stroke: copy of every path's staged output to its final place.
struct StrokeCompact {
    const uint32_t* path_nv;     // exact per-path counts (vertices, indices)
    const uint32_t* path_ni;
    const uint32_t* stage_voff;  // staging offsets (n + 1): the capacities' prefix sums
    const uint32_t* stage_ioff;
    const float* stage_v;
    const uint32_t* stage_i;     // path-local vertex ids
    const uint32_t* vtx_off;     // final offsets (n + 1)
    const uint32_t* idx_off;
    float* out_vertices;
    uint32_t* out_indices;
    uint32_t vertex_offset;
    uint32_t local_ids;
    uint32_t n_paths;
    uint32_t* fallback;          // paths that outgrew their staging room
    uint32_t* n_fallback;
    // split walks under graph replay: the number of items is on the device, and items
    // that would land beyond the output arrays are skipped and flagged (err bit 2)
    const unsigned long long* n_extra_dev;
    uint32_t cap_extra;
    uint64_t out_vcap, out_icap;
    uint32_t* err;
};

namespace fillk {
// Capacities of one wave's fill working arrays (events, heap, active edges, spans,
// monotone nodes, pending edges).
struct FillCaps {
    uint32_t ev, heap, act, span, node, below;
};
}  // namespace fillk

struct FillLaunch {
    PathBatchDev in;
    const lyg_fill_options* opts;  // per path (opts_stride 1) or shared (stride 0)
    uint32_t opts_stride;
    uint8_t* work;                 // per-wave working memory (wave slot * work_stride)
    uint64_t work_stride;
    fillk::FillCaps caps;          // capacities of the per-wave arrays
    uint32_t hot_bytes;            // per-wave dynamic LDS for the hot arrays (0: none)
    uint32_t waves_per_block;
    // placement form of the hot arrays: 0 they start in LDS and move to global memory when
    // they outgrow it; 1 all in LDS; 2 the arrays touched on every event in LDS, spans and
    // nodes in global memory.  Forms 1 and 2 use the lds_caps capacities (act, span, node,
    // heap, below) and stop a path that outgrows one with the retry status.
    uint32_t form;
    fillk::FillCaps lds_caps;
    float* stage_v;                // per-path staging, path i at stage_v_off[i] (vertices)
    uint32_t* stage_t;             // and stage_t_off[i] (triangles)
    const uint64_t* stage_v_off;
    const uint64_t* stage_t_off;
    uint32_t* path_nv;
    uint32_t* path_ni;
    uint32_t* path_status;
    uint32_t* path_round;          // the round whose staging holds the path's output
    uint32_t* path_hint;           // retry status: estimated vertices the path needs
    uint32_t round;
    const uint32_t* path_list;     // paths to process (nullptr: 0..n_list)
    uint32_t n_list;
    uint32_t* work_counter;
    // copy pass
    const uint32_t* vtx_off;
    const uint32_t* idx_off;
    float* out_vertices;
    uint32_t* out_indices;
    uint32_t vertex_offset;
    uint32_t local_ids;            // ids count from vertex_offset in every path
    // output capacity (vertices, indices): a path whose range ends past it is not copied
    uint64_t vertex_capacity;
    uint64_t index_capacity;
};

// Variable-width stroke (stroke_vw_kernel): one lane per path, count then emit.
struct StrokeVWLaunch {
    PathBatchDev in;
    const lyg_stroke_options* opts;
    uint32_t opts_stride;
    const float* widths;     // per point of the batch: the width factor of its endpoint
    uint32_t widths_stride;  // floats per point (custom attributes per endpoint)
    uint32_t* path_nv;       // count pass
    uint32_t* path_ni;
    uint32_t* path_status;
    const uint32_t* vtx_off;  // emit pass
    const uint32_t* idx_off;
    float* out_vertices;
    uint32_t* out_indices;
    uint32_t vertex_offset;
    uint32_t* err;
    uint32_t paths_per_wave;  // lanes of a wave that take paths (1-64); the rest idle
    // optional: every vertex's interpolated custom attributes (n_attrs per vertex) from the
    // attribute table (n_attrs per point of the batch)
    const float* attrs;
    uint32_t n_attrs;
    float* out_attrs;
};

struct FlattenLaunch {
    const int32_t* kinds;  // 0 line, 1 quadratic, 2 cubic
    const float* ctrl;     // 8 floats per segment: from, ctrl1, ctrl2, to
    uint32_t n;
    float tolerance;
    uint32_t* counts;         // count pass
    const uint32_t* offsets;  // emit pass (n + 1)
    float* out_points;
    uint32_t* runs;           // per run of 16 output points: segment, t bits (null: the emit
                              // searches the offsets and computes t)
    float* polys;             // with runs: 12 floats per segment (polynomial form, step, n)
};

// Path validators (lyon_algorithms hit_test.rs / area.rs) over the batched layout.
struct ValidateLaunch {
    const uint8_t* verbs;
    const float* points;
    const uint32_t* verb_off;   // n_paths + 1
    const uint32_t* point_off;  // n_paths + 1
    uint32_t n_paths;
    float tolerance;
    // winding numbers: one query point per entry, against path query_path[i]
    const float* queries;  // x, y pairs
    const uint32_t* query_path;
    uint32_t n_queries;
    int fill_rule;          // 0 EvenOdd, 1 NonZero (for `inside`)
    int32_t* winding;       // n_queries
    uint8_t* inside;        // n_queries or null
    // areas: one per path
    float* areas;
};

// SVG path data parse (svg.hip): one lane per string.
struct SvgLaunch {
    const uint8_t* text;
    const uint32_t* text_off;  // n + 1
    uint32_t n;
    int num_attributes;
    int stop_at;               // ASCII character or -1
    uint32_t* counts_v;
    uint32_t* counts_p;
    int32_t* errors;           // 4 per string or null
    uint32_t* n_errors;        // [0] strings with an error, [1] staging rooms overflowed
    uint8_t* verbs;            // staging rooms (see svg_parse_kernel)
    float* points;
    // compaction
    const uint32_t* off_v;
    const uint32_t* off_p;
};

}  // namespace lyg
