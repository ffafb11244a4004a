// ORACLE — test infrastructure only (see refgeom.h).
//
// CPU restatement of lyon_tessellation's FillTessellator
// (crates/tessellation/src/fill.rs, event_queue.rs, monotone.rs).
#pragma once
#include "refgeom.h"
#include "refpath.h"

namespace reffill {
using namespace refgeom;

enum FillRule { EVEN_ODD = 0, NON_ZERO = 1 };

struct FillOptions {
    float tolerance = 0.1f;
    int fill_rule = EVEN_ODD;
    int sweep_orientation = 1;  // 0 horizontal, 1 vertical (lib.rs Orientation order)
    bool handle_intersections = true;
};

// TessellationError mapping used by the oracle's status codes.
enum Status {
    OK = 0,
    TOO_MANY_VERTICES = 1,
    POSITION_IS_NAN = 3,
    TOLERANCE_IS_NAN = 4,
    INTERNAL = 5,
    PANIC = 6,  // the reference panics (Rust index out of bounds / unreachable!())
};

struct Output {
    std::vector<V2>* vertices;
    std::vector<uint32_t>* indices;
    uint32_t vertex_offset = 0;
    size_t first_vertex = 0, first_index = 0;
    void begin_geometry() {
        first_vertex = vertices->size();
        first_index = indices->size();
    }
    bool add_vertex(V2 p, uint32_t* id) {
        uint64_t v = (uint64_t)vertices->size() + vertex_offset;
        if (v > 0xfffffffeull) return false;
        vertices->push_back(p);
        *id = (uint32_t)v;
        return true;
    }
    void add_triangle(uint32_t a, uint32_t b, uint32_t c) {
        indices->push_back(a);
        indices->push_back(b);
        indices->push_back(c);
    }
    void abort_geometry() {
        vertices->resize(first_vertex);
        indices->resize(first_index);
    }
};

int fill_path(const refpath::PathView& path, const FillOptions& opts, Output& out);

// The flattened edges of the event builder (x0 y0 x1 y1 each, in path direction).
void flattened_edges(const refpath::PathView& path, float tolerance, std::vector<float>& out);

}  // namespace reffill
