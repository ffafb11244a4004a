// ORACLE — test infrastructure only (see refgeom.h).
//
// lyon_path's Path storage (crates/path/src/path.rs:29-36, 75-79, 516-611) and its
// event iterator Iter::next (path.rs:843-909), restated over plain arrays.
#pragma once
#include <cmath>
#include <cstdint>

#include "refgeom.h"

namespace refpath {
using refgeom::V2;

// Verb order of path.rs:29-36.
enum Verb : uint8_t { LINE = 0, QUAD = 1, CUBIC = 2, BEGIN = 3, CLOSE = 4, END = 5 };

struct PathView {
    const uint8_t* verbs;
    uint32_t n_verbs;
    const float* points;  // x, y pairs
    uint32_t n_points;
};

enum EventKind { EV_BEGIN, EV_LINE, EV_QUAD, EV_CUBIC, EV_END };
struct Event {
    EventKind kind;
    V2 from, ctrl1, ctrl2, to;  // Begin: to = at; End: from = last, to = first
    bool close;
};

// Iter (path.rs:817-909).  Reading past the end of points yields NaN (PointIter::next).
struct Iter {
    const PathView& p;
    uint32_t vi = 0, pi = 0;
    V2 current{0, 0}, first{0, 0};
    explicit Iter(const PathView& path) : p(path) {}
    V2 next_point() {
        if (pi >= p.n_points) return V2{NAN, NAN};
        V2 r{p.points[2 * pi], p.points[2 * pi + 1]};
        ++pi;
        return r;
    }
    bool next(Event& e) {
        if (vi >= p.n_verbs) return false;
        uint8_t v = p.verbs[vi++];
        switch (v) {
            case BEGIN:
                current = next_point();
                first = current;
                e.kind = EV_BEGIN;
                e.to = current;
                return true;
            case LINE:
                e.kind = EV_LINE;
                e.from = current;
                current = next_point();
                e.to = current;
                return true;
            case QUAD:
                e.kind = EV_QUAD;
                e.from = current;
                e.ctrl1 = next_point();
                current = next_point();
                e.to = current;
                return true;
            case CUBIC:
                e.kind = EV_CUBIC;
                e.from = current;
                e.ctrl1 = next_point();
                e.ctrl2 = next_point();
                current = next_point();
                e.to = current;
                return true;
            case CLOSE:
                e.kind = EV_END;
                e.from = current;
                (void)next_point();
                e.to = first;
                e.close = true;
                return true;
            case END:
            default:
                e.kind = EV_END;
                e.from = current;
                current = first;
                e.to = first;
                e.close = false;
                return true;
        }
    }
};

}  // namespace refpath
