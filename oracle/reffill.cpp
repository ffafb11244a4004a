// ORACLE — test infrastructure only (see refgeom.h).
//
// FillTessellator restated from crates/tessellation/src/fill.rs (sweep line),
// event_queue.rs (event queue and curve flattening) and monotone.rs (monotone
// polygon triangulation).  Function names follow the reference; each cites its lines.
#include "reffill.h"

#include <algorithm>
#include <cfloat>
#include <memory>
#include <utility>

namespace reffill {

static const uint32_t INVALID_EVENT_ID = 0xffffffffu;
static const uint32_t INVALID_VERTEX = 0xffffffffu;

// fill.rs:2122-2151
static int compare_positions(V2 a, V2 b) {
    if (a.y > b.y) return 1;
    if (a.y < b.y) return -1;
    if (a.x > b.x) return 1;
    if (a.x < b.x) return -1;
    return 0;
}
static bool is_after(V2 a, V2 b) { return a.y > b.y || (a.y == b.y && a.x > b.x); }
static bool is_near(V2 a, V2 b) { return square_length(a - b) < 0.000000001f; }
static bool points_are_equal(V2 a, V2 b) { return a == b; }
// fill.rs:54-66
static float fmax_(float a, float b) { return a > b ? a : b; }
static float slope(V2 v) { return v.x / fmaxr(v.y, FLT_MIN); }
// path/src/lib.rs:267-280
static bool is_in(int rule, int16_t w) { return rule == EVEN_ODD ? (w % 2 != 0) : (w != 0); }

// ---------------------------------------------------------------------------------
// Event queue (event_queue.rs:22-450)
// ---------------------------------------------------------------------------------
struct Event {
    uint32_t next_sibling, next_event;
    V2 position;
};
struct EdgeData {
    V2 to;
    float range_start, range_end;
    int16_t winding;
    bool is_edge;
};

struct EventQueue {
    std::vector<Event> events;
    std::vector<EdgeData> edge_data;
    uint32_t first = INVALID_EVENT_ID;
    bool sorted = false;

    void push_unsorted(V2 p) { events.push_back(Event{INVALID_EVENT_ID, INVALID_EVENT_ID, p}); }
    bool valid_id(uint32_t id) const { return id != INVALID_EVENT_ID; }
    uint32_t next_id(uint32_t id) const { return events[id].next_event; }
    uint32_t next_sibling_id(uint32_t id) const { return events[id].next_sibling; }
    V2 position(uint32_t id) const { return events[id].position; }

    // :252-274
    void insert_into_sorted_list(uint32_t idx, V2 position, uint32_t after) {
        uint32_t prev = after, current = after;
        while (valid_id(current)) {
            V2 pos = events[current].position;
            if (pos == position) {
                events[idx].next_sibling = events[current].next_sibling;
                events[current].next_sibling = idx;
                return;
            } else if (is_after(pos, position)) {
                events[prev].next_event = idx;
                events[idx].next_event = current;
                return;
            }
            prev = current;
            current = next_id(current);
        }
        events[prev].next_event = idx;
    }
    // :188-213
    uint32_t insert_sorted(V2 position, const EdgeData& data, uint32_t after) {
        uint32_t idx = (uint32_t)events.size();
        events.push_back(Event{INVALID_EVENT_ID, INVALID_EVENT_ID, position});
        edge_data.push_back(data);
        insert_into_sorted_list(idx, position, after);
        return idx;
    }
    // :215-228
    void insert_sibling(uint32_t sibling, V2 position, const EdgeData& data) {
        uint32_t idx = (uint32_t)events.size();
        uint32_t ns = events[sibling].next_sibling;
        events.push_back(Event{ns, INVALID_EVENT_ID, position});
        edge_data.push_back(data);
        events[sibling].next_sibling = idx;
    }
    // :230-250
    void vertex_event_sorted(V2 position, uint32_t after) {
        uint32_t idx = (uint32_t)events.size();
        push_unsorted(position);
        edge_data.push_back(EdgeData{V2{NAN, NAN}, 0.0f, 0.0f, 0, false});
        insert_into_sorted_list(idx, position, after);
    }
    // :301-410
    void sort() {
        sorted = true;
        if (events.empty()) return;
        first = merge_sort(0, events.size());
    }
    uint32_t merge_sort(size_t a, size_t b) {
        size_t split = (a + b) / 2;
        if (split == a) return (uint32_t)a;
        uint32_t ah = merge_sort(a, split);
        uint32_t bh = merge_sort(split, b);
        return merge(ah, bh);
    }
    uint32_t find_last_sibling(uint32_t id) const {
        uint32_t cur = id;
        while (valid_id(events[cur].next_sibling)) cur = events[cur].next_sibling;
        return cur;
    }
    uint32_t merge(uint32_t a, uint32_t b) {
        if (a == INVALID_EVENT_ID) return b;
        if (b == INVALID_EVENT_ID) return a;
        bool first_ = true;
        uint32_t head = INVALID_EVENT_ID, prev = INVALID_EVENT_ID;
        for (;;) {
            if (a == INVALID_EVENT_ID) {
                if (!first_) events[prev].next_event = b;
                break;
            }
            if (b == INVALID_EVENT_ID) {
                if (!first_) events[prev].next_event = a;
                break;
            }
            uint32_t node;
            int c = compare_positions(events[a].position, events[b].position);
            if (c < 0) {
                node = a;
                a = events[a].next_event;
            } else if (c > 0) {
                node = b;
                b = events[b].next_event;
            } else {
                uint32_t as = find_last_sibling(a);
                events[as].next_sibling = b;
                b = events[b].next_event;
                continue;
            }
            if (first_) {
                first_ = false;
                head = node;
            } else {
                events[prev].next_event = node;
            }
            prev = node;
        }
        if (head == INVALID_EVENT_ID) head = a;
        return head;
    }
};

// EventQueueBuilder (event_queue.rs:451-905), vertical orientation only at entry;
// horizontal orientation is applied by the caller through reorient().
struct EventQueueBuilder {
    V2 current{NAN, NAN}, prev{NAN, NAN}, second{NAN, NAN};
    uint32_t nth = 0;
    EventQueue queue;
    float tolerance;

    void vertex_event(V2 at) {
        queue.push_unsorted(at);
        queue.edge_data.push_back(EdgeData{V2{NAN, NAN}, 0.0f, 0.0f, 0, false});
    }
    void vertex_event_on_curve(V2 at, float t) {
        queue.push_unsorted(at);
        queue.edge_data.push_back(EdgeData{V2{NAN, NAN}, t, t, 0, false});
    }
    // :661-682
    void end(V2 first) {
        if (nth == 0) return;
        line_segment(first, 0.0f, 1.0f);
        if (is_after(first, prev) && is_after(first, second)) vertex_event(first);
        nth = 0;
    }
    // :684-690
    void begin(V2 to) {
        nth = 0;
        current = to;
    }
    // :693-726
    void add_edge(V2 from, V2 to, int16_t winding, float t0, float t1) {
        if (from == to) return;
        V2 evt_pos = from, evt_to = to;
        if (is_after(evt_pos, to)) {
            evt_to = evt_pos;
            evt_pos = to;
            std::swap(t0, t1);
            winding = (int16_t)(winding * -1);
        }
        queue.push_unsorted(evt_pos);
        queue.edge_data.push_back(EdgeData{evt_to, t0, t1, winding, true});
        nth += 1;
    }
    // :728-756
    void line_segment(V2 to, float t0, float t1) {
        V2 from = current;
        if (from == to) return;
        if (is_after(from, to) && nth > 0 && is_after(from, prev)) vertex_event(from);
        if (nth == 0) second = to;
        add_edge(from, to, 1, t0, t1);
        prev = current;
        current = to;
    }
    // :758-824 / :826-900 (shared tail)
    template <class Flatten>
    void curve_segment(V2 orig_from, V2 orig_to, bool needs_swap, Flatten&& flatten) {
        int16_t winding = needs_swap ? -1 : 1;
        V2 prev_ = needs_swap ? orig_to : orig_from;
        bool have_first = false;
        V2 first{0, 0};
        bool is_first_edge = nth == 0;
        flatten([&](V2 lf, V2 lt, float ts, float te) {
            if (lf == lt) return;
            if (!have_first) {
                have_first = true;
                first = lt;
            } else if (is_after(lf, lt) && is_after(lf, prev_)) {
                vertex_event_on_curve(lf, ts);
            }
            add_edge(lf, lt, winding, ts, te);
            prev_ = lf;
        });
        if (have_first) {
            V2 second_ = needs_swap ? prev_ : first;
            V2 previous = needs_swap ? first : prev_;
            if (is_first_edge) {
                second = second_;
            } else if (is_after(orig_from, prev) && is_after(orig_from, second_)) {
                vertex_event(orig_from);
            }
            prev = previous;
            current = orig_to;
        }
    }
    void quadratic_bezier_segment(V2 ctrl, V2 to) {
        V2 from = current;
        bool needs_swap = is_after(from, to);
        V2 f = from, t = to;
        if (needs_swap) std::swap(f, t);
        float tol = tolerance;
        curve_segment(from, to, needs_swap, [&](auto&& cb) {
            quad_flatten_with_t(f, ctrl, t, tol, cb);
        });
    }
    void cubic_bezier_segment(V2 c1, V2 c2, V2 to) {
        V2 from = current;
        bool needs_swap = is_after(from, to);
        V2 f = from, a = c1, b = c2, t = to;
        if (needs_swap) {
            std::swap(f, t);
            std::swap(a, b);
        }
        float tol = tolerance;
        curve_segment(from, to, needs_swap, [&](auto&& cb) {
            cubic_flatten_with_t(f, a, b, t, tol, cb);
        });
    }
};

// ---------------------------------------------------------------------------------
// Monotone tessellators (monotone.rs)
// ---------------------------------------------------------------------------------
enum Side { LEFT = 0, RIGHT = 1 };
struct MonotoneVertex {
    V2 pos;
    uint32_t id;
    int side;
};
struct Tri {
    uint32_t a, b, c;
};

// BasicMonotoneTessellator (monotone.rs:9-139)
struct BasicMono {
    std::vector<MonotoneVertex> stack;
    MonotoneVertex previous{V2{0, 0}, 0, LEFT};
    std::vector<Tri> triangles;

    void begin(V2 pos, uint32_t id) {
        MonotoneVertex first{pos, id, LEFT};
        previous = first;
        triangles.clear();
        stack.clear();
        stack.push_back(first);
    }
    void push_triangle_ids(uint32_t a, uint32_t b, uint32_t c) { triangles.push_back(Tri{a, b, c}); }
    void push_triangle(const MonotoneVertex& a, const MonotoneVertex& b, const MonotoneVertex& c) {
        push_triangle_ids(a.id, b.id, c.id);
    }
    void monotone_vertex(MonotoneVertex current) {
        bool changed_side = current.side != previous.side;
        if (changed_side) {
            for (size_t i = 0; i + 1 < stack.size(); ++i) {
                MonotoneVertex a = stack[i], b = stack[i + 1];
                bool winding = cross(a.pos - b.pos, current.pos - b.pos) >= 0.0f;
                if (!winding) std::swap(a, b);
                push_triangle(a, b, current);
            }
            stack.clear();
            stack.push_back(previous);
        } else {
            bool has_last = !stack.empty();
            MonotoneVertex last_popped{};
            if (has_last) {
                last_popped = stack.back();
                stack.pop_back();
            }
            while (!stack.empty()) {
                MonotoneVertex a = last_popped;
                MonotoneVertex b = stack.back();
                if (current.side == RIGHT) std::swap(a, b);
                float c = cross(current.pos - b.pos, a.pos - b.pos);
                if (c >= 0.0f) {
                    push_triangle(b, a, current);
                    has_last = !stack.empty();
                    if (has_last) {
                        last_popped = stack.back();
                        stack.pop_back();
                    }
                } else {
                    break;
                }
            }
            if (has_last) stack.push_back(last_popped);
        }
        stack.push_back(current);
        previous = current;
    }
    void vertex(V2 pos, uint32_t id, int side) { monotone_vertex(MonotoneVertex{pos, id, side}); }
    void end(V2 pos, uint32_t id) {
        int side = 1 - previous.side;
        vertex(pos, id, side);
        stack.clear();
    }
    void flush(Output& out) {
        for (const Tri& t : triangles) out.add_triangle(t.a, t.b, t.c);
        triangles.clear();
    }
};

// SideEvents + AdvancedMonotoneTessellator (monotone.rs:189-404)
struct SideEvents {
    std::vector<uint32_t> events;
    V2 reference_point{0, 0};
    float conservative_reference_x = 0;
    V2 prev{0, 0};
    MonotoneVertex last{V2{0, 0}, 0, LEFT};
    void push(MonotoneVertex v) {
        events.push_back(v.id);
        prev = last.pos;
        last = v;
    }
};

static bool flush_side(SideEvents& side, int s, BasicMono& tess, MonotoneVertex* out) {
    size_t len = side.events.size();
    if (len < 2) return false;
    size_t step = 1;
    while (step * 2 < len) {
        size_t last_index = 0;
        size_t imax = (len - 1) / (2 * step);
        for (size_t i = 0; i < imax; ++i) {
            size_t a = i * 2 * step;
            size_t b = a + step;
            last_index = b + step;
            if (s == RIGHT) std::swap(a, b);
            tess.push_triangle_ids(side.events[a], side.events[b], side.events[last_index]);
        }
        if (last_index + step < len) {
            size_t b = last_index, c = last_index + step;
            if (s == RIGHT) std::swap(b, c);
            tess.push_triangle_ids(side.events[0], side.events[b], side.events[c]);
        }
        step *= 2;
    }
    side.events.clear();
    side.push(side.last);
    side.reference_point = side.last.pos;
    *out = side.last;
    return true;
}

struct AdvancedMono {
    BasicMono tess;
    SideEvents left, right;
    bool flushing = false;

    void begin(V2 pos, uint32_t id) {
        tess.begin(pos, id);
        left.reference_point = pos;
        left.conservative_reference_x = pos.x;
        right.reference_point = pos;
        right.conservative_reference_x = pos.x;
        left.prev = pos;
        right.prev = pos;
        flushing = false;
        left.events.clear();
        right.events.clear();
        left.push(MonotoneVertex{pos, id, LEFT});
        right.push(MonotoneVertex{pos, id, RIGHT});
    }
    void vertex(V2 pos, uint32_t id, int side) {
        if (side == LEFT) {
            left.reference_point.x = fmaxr(left.reference_point.x, pos.x);
            left.conservative_reference_x =
                fmaxr(left.conservative_reference_x, left.reference_point.x);
        } else {
            right.reference_point.x = fminr(right.reference_point.x, pos.x);
            right.conservative_reference_x =
                fminr(right.conservative_reference_x, right.reference_point.x);
        }
        float dx = right.conservative_reference_x - left.conservative_reference_x;
        SideEvents& side_ev = side == LEFT ? left : right;
        SideEvents& opp = side == LEFT ? right : left;
        float dy = pos.y - side_ev.reference_point.y;
        bool sides_are_close = dx < dy * 0.1f;
        size_t len = side_ev.events.size();
        bool outward_turn = false;
        if (!sides_are_close && len >= 2) {
            float sign = side == LEFT ? 1.0f : -1.0f;
            V2 prev = side_ev.prev;
            V2 last = side_ev.last.pos;
            outward_turn = cross(prev - last, pos - last) * sign < 0.0f;
        }
        if (outward_turn || sides_are_close) {
            bool must_flush_opposite = is_after(side_ev.last.pos, opp.last.pos);
            MonotoneVertex v;
            if (must_flush_opposite) {
                if (flush_side(opp, 1 - side, tess, &v)) {
                    tess.monotone_vertex(v);
                    side_ev.conservative_reference_x = side_ev.reference_point.x;
                }
            }
            if (flush_side(side_ev, side, tess, &v)) {
                tess.monotone_vertex(v);
                opp.conservative_reference_x = opp.reference_point.x;
            }
        }
        side_ev.push(MonotoneVertex{pos, id, side});
    }
    void end(V2 pos, uint32_t id) {
        MonotoneVertex a, b;
        bool ha = flush_side(left, LEFT, tess, &a);
        bool hb = flush_side(right, RIGHT, tess, &b);
        if (ha && !hb) {
            tess.monotone_vertex(a);
        } else if (!ha && hb) {
            tess.monotone_vertex(b);
        } else if (ha && hb) {
            if (is_after(a.pos, b.pos)) std::swap(a, b);
            tess.monotone_vertex(a);
            tess.monotone_vertex(b);
        }
        tess.end(pos, id);
    }
    void flush(Output& out) { tess.flush(out); }
};

// ---------------------------------------------------------------------------------
// Sweep (fill.rs)
// ---------------------------------------------------------------------------------
struct WindingState {
    int32_t span_index = -1;
    int16_t number = 0;
    bool in = false;
    void update(int rule, int16_t w) {
        number = (int16_t)(number + w);
        in = is_in(rule, number);
        if (in) span_index += 1;
    }
};

struct ActiveEdge {
    V2 from, to;
    int16_t winding;
    bool is_merge;
    uint32_t from_id;
    uint32_t src_edge;
    float range_end;
    float min_x() const { return fminr(from.x, to.x); }
    float max_x() const { return fmax_(from.x, to.x); }
    // fill.rs:181-195 with LineSegment::solve_x_for_y (line.rs:56-71)
    float solve_x_for_y(float y) const {
        float dy = to.y - from.y;
        float t = dy == 0.0f ? 0.0f : (y - from.y) / dy;
        float x = from.x * (1.0f - t) + to.x * t;
        return fminr(fmaxr(x, min_x()), max_x());
    }
};

struct PendingEdge {
    V2 to;
    float sort_key;
    uint32_t src_edge;
    int16_t winding;
    float range_end;
};

struct ScanResult {
    std::vector<std::pair<int32_t, int>> vertex_events;
    std::vector<size_t> edges_to_split;
    std::vector<int32_t> spans_to_end;
    bool merge_event = false, split_event = false, merge_split_event = false;
    size_t above_start = 0, above_end = 0;
    WindingState winding_before_point;
    void reset() {
        vertex_events.clear();
        edges_to_split.clear();
        spans_to_end.clear();
        merge_event = split_event = merge_split_event = false;
        above_start = above_end = 0;
        winding_before_point = WindingState();
    }
};

enum InternalErr { IE_NONE = 0, IE_ORDER, IE_MERGE_OUTSIDE, IE_SPANS };

// fill.rs:2327-2337
static float remap_t_in_range(float val, float rs, float re) {
    if (re > rs) {
        float d = re - rs;
        return rs + val * d;
    }
    float d = rs - re;
    return re + (1.0f - val) * d;
}

struct Span {
    std::unique_ptr<AdvancedMono> tess;
};

// A Rust panic of the reference (an out-of-bounds index, or unreachable!() on a span that
// has already ended): the oracle reports it as the PANIC status instead of a result.
struct LyonPanic {};

struct Tessellator {
    V2 current_position{-FLT_MAX, -FLT_MAX};
    uint32_t current_vertex = INVALID_VERTEX;
    uint32_t current_event_id = INVALID_EVENT_ID;
    std::vector<ActiveEdge> active;
    std::vector<PendingEdge> edges_below;
    int fill_rule = EVEN_ODD;
    int orientation = 1;
    float tolerance = 0.1f;
    bool assume_no_intersection = false;
    std::vector<Span> spans;
    std::vector<std::unique_ptr<AdvancedMono>> pool;
    ScanResult scan;
    EventQueue events;
    Output* out = nullptr;

    // Spans (fill.rs:223-300)
    // Spans::begin_span / end_span (fill.rs:233-263): Vec::insert and indexing panic
    // out of range, end_span on an ended span reaches unreachable!().
    AdvancedMono& live(int32_t idx) {
        if (idx < 0 || (size_t)idx >= spans.size() || !spans[idx].tess) throw LyonPanic();
        return *spans[idx].tess;
    }
    ActiveEdge& act(size_t i) {
        if (i >= active.size()) throw LyonPanic();
        return active[i];
    }
    void begin_span(int32_t idx, V2 pos, uint32_t vertex) {
        if (idx < 0 || (size_t)idx > spans.size()) throw LyonPanic();
        std::unique_ptr<AdvancedMono> t;
        if (!pool.empty()) {
            t = std::move(pool.back());
            pool.pop_back();
        } else {
            t.reset(new AdvancedMono());
        }
        t->begin(pos, vertex);
        spans.insert(spans.begin() + idx, Span{std::move(t)});
    }
    void end_span(int32_t idx, V2 pos, uint32_t id) {
        live(idx);
        Span& s = spans[idx];
        if (s.tess) {
            s.tess->end(pos, id);
            s.tess->flush(*out);
            pool.push_back(std::move(s.tess));
            s.tess.reset();
        }
    }
    void cleanup_spans() {
        spans.erase(std::remove_if(spans.begin(), spans.end(),
                                   [](const Span& s) { return !s.tess; }),
                    spans.end());
    }

    // fill.rs:869-925
    int initialize_events() {
        current_position = events.position(current_event_id);
        if (std::isnan(current_position.x) || std::isnan(current_position.y))
            return POSITION_IS_NAN;
        V2 position = orientation == 1 ? current_position
                                       : V2{current_position.y, -current_position.x};
        if (!out->add_vertex(position, &current_vertex)) return TOO_MANY_VERTICES;
        uint32_t sib = current_event_id;
        while (events.valid_id(sib)) {
            const EdgeData& e = events.edge_data[sib];
            if (e.is_edge) {
                V2 to = e.to;
                edges_below.push_back(
                    PendingEdge{to, slope(to - current_position), sib, e.winding, e.range_end});
            }
            sib = events.next_sibling_id(sib);
        }
        return OK;
    }

    // fill.rs:1328-1378
    InternalErr is_edge_connecting(const ActiveEdge& e, size_t idx, bool* connecting) {
        if (points_are_equal(current_position, e.to)) {
            *connecting = true;
            return IE_NONE;
        }
        float cx = current_position.x;
        float thr = tolerance;
        float mnx = e.min_x(), mxx = e.max_x();
        if (mxx + thr < cx || e.to.y < current_position.y) return IE_ORDER;
        if (mnx > cx) {
            *connecting = false;
            return IE_NONE;
        }
        float ex;
        if (e.from.y != e.to.y) ex = e.solve_x_for_y(current_position.y);
        else if (mxx >= cx && mnx <= cx) ex = cx;
        else ex = e.to.x;
        if (std::fabs(ex - cx) <= thr) {
            scan.edges_to_split.push_back(idx);
            *connecting = true;
            return IE_NONE;
        }
        if (ex < cx) return IE_ORDER;
        *connecting = false;
        return IE_NONE;
    }

    // fill.rs:1295-1324
    InternalErr check_remaining_edges(size_t idx, float cx) {
        for (size_t i = idx; i < active.size(); ++i) {
            const ActiveEdge& e = active[i];
            if (e.is_merge) continue;
            if (e.max_x() < cx) return IE_ORDER;
            if (points_are_equal(current_position, e.to)) return IE_ORDER;
            if (e.min_x() < cx && e.solve_x_for_y(current_position.y) < cx) return IE_ORDER;
        }
        return IE_NONE;
    }

    // fill.rs:1045-1291
    InternalErr scan_active_edges() {
        scan.reset();
        float cx = current_position.x;
        bool connecting_edges = false;
        size_t idx = 0;
        WindingState winding;
        bool previous_was_merge = false;
        for (const ActiveEdge& e : active) {
            if (e.is_merge) {
                winding.span_index += 1;
                idx += 1;
                previous_was_merge = true;
                continue;
            }
            bool before;
            if (points_are_equal(current_position, e.to)) {
                connecting_edges = true;
                before = false;
            } else if (e.max_x() < cx) {
                before = true;
            } else if (e.min_x() > cx) {
                before = false;
            } else if (e.from.y == e.to.y) {
                connecting_edges = true;
                before = false;
            } else {
                float ex = e.solve_x_for_y(current_position.y);
                if (std::fabs(ex - cx) <= tolerance) {
                    connecting_edges = true;
                    before = false;
                } else if (ex > cx) {
                    before = false;
                } else {
                    before = true;
                }
            }
            if (!before) break;
            winding.update(fill_rule, e.winding);
            previous_was_merge = false;
            idx += 1;
        }
        scan.above_start = idx;
        scan.winding_before_point = winding;
        if (previous_was_merge) {
            scan.winding_before_point.span_index -= 1;
            scan.above_start -= 1;
            if (!connecting_edges) {
                scan.vertex_events.push_back({winding.span_index - 1, RIGHT});
                scan.vertex_events.push_back({winding.span_index, LEFT});
                scan.merge_split_event = true;
            }
        }
        scan.split_event = !connecting_edges && winding.in && !scan.merge_split_event;
        if (connecting_edges) {
            bool in_before = winding.in;
            bool first_connecting_edge = !previous_was_merge;
            for (size_t i = idx; i < active.size(); ++i) {
                const ActiveEdge& e = active[i];
                if (e.is_merge) {
                    if (!winding.in) return IE_MERGE_OUTSIDE;
                    scan.spans_to_end.push_back(winding.span_index);
                    winding.span_index += 1;
                    idx += 1;
                    first_connecting_edge = false;
                    continue;
                }
                bool conn;
                InternalErr err = is_edge_connecting(e, idx, &conn);
                if (err != IE_NONE) return err;
                if (!conn) break;
                if (!first_connecting_edge && winding.in) scan.spans_to_end.push_back(winding.span_index);
                winding.update(fill_rule, e.winding);
                if (winding.in && winding.span_index >= (int32_t)spans.size()) return IE_SPANS;
                idx += 1;
                first_connecting_edge = false;
            }
            bool in_after = winding.in;
            bool merge = in_before && in_after && edges_below.empty() && scan.edges_to_split.empty();
            if (merge) scan.merge_event = true;
            if (in_before) scan.vertex_events.push_back({scan.winding_before_point.span_index, RIGHT});
            if (in_after) scan.vertex_events.push_back({winding.span_index, LEFT});
        }
        scan.above_end = idx;
        return check_remaining_edges(idx, cx);
    }

    // fill.rs:1381-1452
    void process_edges_above() {
        for (auto& ve : scan.vertex_events)
            live(ve.first).vertex(current_position, current_vertex, ve.second);
        for (int32_t si : scan.spans_to_end) end_span(si, current_position, current_vertex);
        cleanup_spans();
        for (size_t ei : scan.edges_to_split) {
            ActiveEdge& e = act(ei);
            V2 to = e.to;
            edges_below.push_back(
                PendingEdge{to, slope(to - current_position), e.src_edge, e.winding, e.range_end});
            e.to = current_position;
        }
        if (scan.merge_event) {
            ActiveEdge& e = act(scan.above_start);
            e.is_merge = true;
            e.from = e.to;
            e.winding = 0;
            e.from_id = current_vertex;
            scan.above_start += 1;
        }
    }

    // fill.rs:2010-2013
    void sort_edges_below() {
        std::sort(edges_below.begin(), edges_below.end(),
                  [](const PendingEdge& a, const PendingEdge& b) { return a.sort_key < b.sort_key; });
    }

    // fill.rs:2046-2106
    void merge_coincident_edges(size_t a_idx, size_t b_idx) {
        V2 a_to = edges_below[a_idx].to, b_to = edges_below[b_idx].to;
        size_t lower, upper;
        bool split;
        int c = compare_positions(a_to, b_to);
        if (c > 0) {
            lower = a_idx;
            upper = b_idx;
            split = true;
        } else if (c < 0) {
            lower = b_idx;
            upper = a_idx;
            split = true;
        } else {
            lower = a_idx;
            upper = b_idx;
            split = false;
        }
        edges_below[upper].winding = (int16_t)(edges_below[upper].winding + edges_below[lower].winding);
        V2 split_point = edges_below[upper].to;
        PendingEdge edge = edges_below[lower];
        edges_below.erase(edges_below.begin() + lower);
        if (!split) return;
        EdgeData src = events.edge_data[edge.src_edge];
        float dy = edge.to.y - current_position.y;
        float t = dy == 0.0f ? 0.0f : (split_point.y - current_position.y) / dy;
        float t_remapped = remap_t_in_range(t, src.range_start, edge.range_end);
        EdgeData ed{edge.to, t_remapped, edge.range_end, edge.winding, true};
        events.insert_sorted(split_point, ed, current_event_id);
    }

    // fill.rs:2016-2043
    void handle_coincident_edges_below() {
        if (edges_below.size() < 2) return;
        for (size_t k = edges_below.size() - 1; k-- > 0;) {
            size_t a_idx = k, b_idx = k + 1;
            float a_slope = edges_below[a_idx].sort_key;
            float b_slope = edges_below[b_idx].sort_key;
            const float THRESHOLD = 0.00005f;
            bool close;
            if (std::fabs(a_slope) <= 1.0f) close = std::fabs(a_slope - b_slope) < THRESHOLD;
            else close = std::fabs(1.0f / a_slope - 1.0f / b_slope) < THRESHOLD;
            if (close) merge_coincident_edges(a_idx, b_idx);
        }
    }

    // fill.rs:1563-1608
    void split_event(size_t left_idx, int32_t left_span) {
        size_t right_idx = left_idx + 1;
        V2 ul = act(left_idx).from, ur = act(right_idx).from;
        int32_t right_span = left_span + 1;
        V2 up;
        uint32_t uid;
        int32_t new_span;
        if (is_after(ul, ur)) {
            up = ul;
            uid = active[left_idx].from_id;
            new_span = left_span;
        } else {
            up = ur;
            uid = active[right_idx].from_id;
            new_span = right_span;
        }
        begin_span(new_span, up, uid);
        live(left_span).vertex(current_position, current_vertex, RIGHT);
        live(right_span).vertex(current_position, current_vertex, LEFT);
    }

    // fill.rs:1455-1524
    void process_edges_below() {
        WindingState winding = scan.winding_before_point;
        sort_edges_below();
        handle_coincident_edges_below();
        if (scan.split_event) split_event(scan.above_start - 1, winding.span_index);
        bool first = true;
        for (const PendingEdge& pe : edges_below) {
            if (!first && winding.in) begin_span(winding.span_index, current_position, current_vertex);
            winding.update(fill_rule, pe.winding);
            first = false;
        }
    }

    // fill.rs:1689-1846
    void process_intersection(double ta, double tb, size_t ai, PendingEdge& eb, D2 bfrom, D2 bto) {
        // LineSegment<f64>::sample = from.lerp(to, tb)
        double one_t = 1.0 - tb;
        V2 ip{(float)(one_t * bfrom.x + tb * bto.x), (float)(one_t * bfrom.y + tb * bto.y)};
        ActiveEdge& ae = active[ai];
        if (current_position == ip) {
            ae.from = ip;
            EdgeData& src = events.edge_data[ae.src_edge];
            src.range_start = remap_t_in_range((float)ta, src.range_start, ae.range_end);
            return;
        }
        if (!is_after(ip, current_position)) ip.y = std::nextafter(current_position.y, INFINITY);
        if (is_near(ip, eb.to)) ip = eb.to;
        else if (is_near(ip, ae.to)) ip = ae.to;
        EdgeData a_src = events.edge_data[ae.src_edge];
        EdgeData b_src = events.edge_data[eb.src_edge];
        bool have_inserted = false;
        uint32_t inserted = 0;
        bool flipped_active = false;
        if (ae.to != ip && ae.from != ip) {
            float rta = remap_t_in_range((float)ta, a_src.range_start, ae.range_end);
            if (is_after(ae.to, ip)) {
                inserted = events.insert_sorted(ip, EdgeData{ae.to, rta, ae.range_end, ae.winding, true},
                                                current_event_id);
                have_inserted = true;
            } else {
                flipped_active = true;
                events.insert_sorted(ae.to, EdgeData{ip, ae.range_end, rta, (int16_t)-ae.winding, true},
                                     current_event_id);
            }
            ae.to = ip;
            ae.range_end = rta;
        }
        if (eb.to != ip && current_position != ip) {
            float rtb = remap_t_in_range((float)tb, b_src.range_start, eb.range_end);
            if (is_after(eb.to, ip)) {
                EdgeData ed{eb.to, rtb, eb.range_end, eb.winding, true};
                if (have_inserted) events.insert_sibling(inserted, ip, ed);
                else events.insert_sorted(ip, ed, current_event_id);
            } else {
                events.insert_sorted(eb.to, EdgeData{ip, eb.range_end, rtb, (int16_t)-eb.winding, true},
                                     current_event_id);
                if (flipped_active) events.vertex_event_sorted(ip, current_event_id);
            }
            eb.to = ip;
            eb.range_end = rtb;
        }
    }

    // LineSegment<f64>::intersection_t (line.rs:217-252)
    static bool intersection_t(D2 af, D2 at, D2 bf, D2 bt, double* ta, double* tb) {
        auto eq = [](D2 p, D2 q) { return p.x == q.x && p.y == q.y; };
        if (eq(at, bt) || eq(af, bf) || eq(af, bt) || eq(at, bf)) return false;
        D2 v1 = at - af, v2 = bt - bf;
        double c = crossd(v1, v2);
        if (c == 0.0) return false;
        double sign = c > 0.0 ? 1.0 : (c < 0.0 ? -1.0 : c);
        double abs_c = std::fabs(c);
        D2 v3 = bf - af;
        double t = crossd(v3, v2) * sign;
        double u = crossd(v3, v1) * sign;
        if (t < 0.0 || t > abs_c || u < 0.0 || u > abs_c) return false;
        *ta = t / abs_c;
        *tb = u / abs_c;
        return true;
    }

    // fill.rs:1611-1686
    void handle_intersections(size_t skip_a, size_t skip_b) {
        std::vector<PendingEdge> eb = std::move(edges_below);
        edges_below.clear();
        for (PendingEdge& e : eb) {
            float bmin = fminr(current_position.x, e.to.x);
            float bmax = fmax_(current_position.x, e.to.x);
            D2 bf = d2(current_position), bt = d2(e.to);
            double tb_min = 1.0;
            bool found = false;
            double fta = 0, ftb = 0;
            size_t fi = 0;
            for (size_t i = 0; i < active.size(); ++i) {
                if (i >= skip_a && i < skip_b) continue;
                const ActiveEdge& ae = active[i];
                if (ae.is_merge || bmin > ae.max_x()) continue;
                if (bmax < ae.min_x()) continue;
                double ta, tb;
                if (intersection_t(d2(ae.from), d2(ae.to), bf, bt, &ta, &tb)) {
                    if (tb < tb_min && tb > 0.0 && ta > 0.0 && ta <= 1.0) {
                        tb_min = tb;
                        found = true;
                        fta = ta;
                        ftb = tb;
                        fi = i;
                    }
                }
            }
            if (found) process_intersection(fta, ftb, fi, e, bf, bt);
        }
        edges_below = std::move(eb);
    }

    // fill.rs:1527-1561
    void update_active_edges() {
        size_t a = scan.above_start, b = scan.above_end;
        if (!assume_no_intersection) handle_intersections(a, b);
        V2 from = current_position;
        uint32_t from_id = current_vertex;
        std::vector<ActiveEdge> repl;
        repl.reserve(edges_below.size());
        for (const PendingEdge& e : edges_below)
            repl.push_back(ActiveEdge{from, e.to, e.winding, false, from_id, e.src_edge, e.range_end});
        edges_below.clear();
        active.erase(active.begin() + a, active.begin() + b);
        active.insert(active.begin() + a, repl.begin(), repl.end());
    }

    // fill.rs:929-965
    InternalErr process_events() {
        InternalErr e = scan_active_edges();
        if (e != IE_NONE) return e;
        process_edges_above();
        process_edges_below();
        update_active_edges();
        return IE_NONE;
    }

    // fill.rs:1848-1953
    void sort_active_edges() {
        float y = current_position.y;
        std::vector<std::pair<float, size_t>> keys;
        keys.reserve(active.size());
        bool has_merge = false;
        float prev_x = NAN;
        for (size_t i = 0; i < active.size(); ++i) {
            const ActiveEdge& e = active[i];
            if (e.is_merge) {
                has_merge = true;
                keys.push_back({prev_x, i});
            } else {
                bool eq_to = e.to.y == y, eq_from = e.from.y == y;
                float x;
                if (eq_to && eq_from) {
                    float cx = current_position.x;
                    x = (e.max_x() >= cx && e.min_x() <= cx) ? current_position.x : e.min_x();
                } else if (eq_from) {
                    x = e.from.x;
                } else if (eq_to) {
                    x = e.to.x;
                } else {
                    x = e.solve_x_for_y(y);
                }
                keys.push_back({fmax_(x, e.min_x()), i});
                prev_x = x;
            }
        }
        // keys.sort_by(|a, b| a.0.partial_cmp(&b.0).unwrap() ..): a NaN key (a merge
        // vertex ahead of every edge) panics as soon as two keys are compared.
        if (keys.size() >= 2)
            for (auto& k : keys)
                if (std::isnan(k.first)) throw LyonPanic();
        std::stable_sort(keys.begin(), keys.end(), [&](const std::pair<float, size_t>& a,
                                                        const std::pair<float, size_t>& b) {
            if (a.first < b.first) return true;
            if (a.first > b.first) return false;
            const ActiveEdge& ea = active[a.second];
            const ActiveEdge& eb = active[b.second];
            if (!ea.is_merge && !eb.is_merge) {
                float sa = slope(ea.to - ea.from), sb = slope(eb.to - eb.from);
                return sb < sa;
            }
            if (!ea.is_merge && eb.is_merge) return true;
            return false;
        });
        std::vector<ActiveEdge> na;
        na.reserve(active.size());
        for (auto& k : keys) na.push_back(active[k.second]);
        active = std::move(na);
        if (!has_merge) return;
        int16_t wn = 0;
        for (size_t i = 0; i < active.size(); ++i) {
            bool needs_swap;
            if (active[i].is_merge) {
                needs_swap = !is_in(fill_rule, wn);
            } else {
                wn = (int16_t)(wn + active[i].winding);
                needs_swap = false;
            }
            if (needs_swap) {
                int16_t w = wn;
                size_t k = i;
                while (k > 0) {
                    w = (int16_t)(w - active[k - 1].winding);
                    std::swap(active[k], active[k - 1]);
                    if (is_in(fill_rule, w)) break;
                    k -= 1;
                }
            }
        }
    }

    // fill.rs:1956-2007
    void recover_from_error() {
        sort_active_edges();
        size_t len = active.size();
        if (len > 1 && active[len - 1].is_merge) std::swap(active[len - 1], active[len - 2]);
        WindingState w;
        for (const ActiveEdge& e : active) {
            if (e.is_merge) w.span_index += 1;
            else w.update(fill_rule, e.winding);
            if (w.span_index >= (int32_t)spans.size()) begin_span(w.span_index, e.from, e.from_id);
        }
        while ((int64_t)spans.size() > (int64_t)w.span_index + 1) {
            if (spans.back().tess) spans.back().tess->flush(*out);
            spans.pop_back();
        }
    }

    // fill.rs:836-867
    int tessellator_loop() {
        current_event_id = events.first;
        while (events.valid_id(current_event_id)) {
            int st = initialize_events();
            if (st != OK) return st;
            if (process_events() != IE_NONE) {
                recover_from_error();
                if (process_events() != IE_NONE) return INTERNAL;
            }
            current_event_id = events.next_id(current_event_id);
        }
        return OK;
    }
};

// The edges the event builder produces for `path` (vertical sweep), each as x0 y0 x1 y1
// in the path's own direction: the polygon a fill covers.
void flattened_edges(const refpath::PathView& path, float tolerance, std::vector<float>& out) {
    EventQueueBuilder b;
    b.tolerance = tolerance;
    refpath::Iter it(path);
    refpath::Event e;
    while (it.next(e)) {
        switch (e.kind) {
            case refpath::EV_BEGIN: b.begin(e.to); break;
            case refpath::EV_LINE: b.line_segment(e.to, 0.0f, 1.0f); break;
            case refpath::EV_QUAD: b.quadratic_bezier_segment(e.ctrl1, e.to); break;
            case refpath::EV_CUBIC: b.cubic_bezier_segment(e.ctrl1, e.ctrl2, e.to); break;
            case refpath::EV_END: b.end(e.to); break;
        }
    }
    for (size_t i = 0; i < b.queue.events.size(); ++i) {
        const EdgeData& d = b.queue.edge_data[i];
        if (!d.is_edge) continue;
        V2 p = b.queue.events[i].position, q = d.to;
        if (d.winding < 0) std::swap(p, q);
        out.insert(out.end(), {p.x, p.y, q.x, q.y});
    }
}

// FillTessellator::tessellate_path -> tessellate -> tessellate_impl (fill.rs:578-821)
int fill_path(const refpath::PathView& path, const FillOptions& opts, Output& out) {
    Tessellator t;
    // EventQueueBuilder::set_path (event_queue.rs:483-551)
    EventQueueBuilder b;
    b.tolerance = opts.tolerance;
    auto ro = [&](V2 p) { return opts.sweep_orientation == 1 ? p : V2{-p.y, p.x}; };
    refpath::Iter it(path);
    refpath::Event e;
    while (it.next(e)) {
        switch (e.kind) {
            case refpath::EV_BEGIN: b.begin(ro(e.to)); break;
            case refpath::EV_LINE: b.line_segment(ro(e.to), 0.0f, 1.0f); break;
            case refpath::EV_QUAD: b.quadratic_bezier_segment(ro(e.ctrl1), ro(e.to)); break;
            case refpath::EV_CUBIC: b.cubic_bezier_segment(ro(e.ctrl1), ro(e.ctrl2), ro(e.to)); break;
            case refpath::EV_END: b.end(ro(e.to)); break;
        }
    }
    b.queue.sort();
    t.events = std::move(b.queue);

    // tessellate_impl
    if (std::isnan(opts.tolerance) || opts.tolerance <= 0.0f) return TOLERANCE_IS_NAN;
    t.fill_rule = opts.fill_rule;
    t.orientation = opts.sweep_orientation;
    t.assume_no_intersection = !opts.handle_intersections;
    t.tolerance = opts.tolerance * 0.5f;
    t.out = &out;
    out.begin_geometry();
    int st;
    try {
        st = t.tessellator_loop();
    } catch (const LyonPanic&) {
        st = PANIC;
    }
    if (st != OK) {
        out.abort_geometry();
        return st;
    }
    if (t.assume_no_intersection && (!t.active.empty() || !t.spans.empty())) {
        out.abort_geometry();
        return INTERNAL;
    }
    for (Span& s : t.spans)
        if (s.tess) s.tess->flush(out);
    return OK;
}

}  // namespace reffill
