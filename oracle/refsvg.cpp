// ORACLE — test infrastructure only (see refgeom.h).  Loaded through ctypes by tests/;
// never by lyon_amd/.
//
// lyon_extra's SVG path parser restated serially: Source (crates/extra/src/parser.rs:48-126),
// PathParser::parse / parse_path / parse_* (parser.rs:154-462), into BuilderImpl's layout
// (crates/path/src/path.rs:479-545).  Arcs: Arc::from_svg_arc (crates/geom/src/arc.rs:58-146),
// SvgArc::is_straight_line (:534-538) and arc_to_quadratic_beziers_with_t (:668-701).
//
// Trigonometry: Rust's f32::cos / sin / tan call the platform's libm, whose last-ulp results
// are not specified (glibc's cosf is faithfully, not always correctly, rounded).  The arcs
// here use the correctly rounded f32 value (evaluated in double, rounded once).  With
// glibc's cosf instead, ~15% of random arc strings move by ~1e-8 of their size, a few
// ill-conditioned arcs by more, and ~0.3% change their number of quadratics (a sweep angle
// on a multiple of pi/4 rounding the other way).
//
// Numbers: Rust's str::parse::<f32> accepts `-? (digits+ (. digits*)? | . digits+)
// ([eE] [+-]? digits+)?` and rounds correctly (half to even); the collected text is checked
// against that grammar here and converted by the C library's strtof, which glibc rounds
// correctly in the default rounding mode.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "refgeom.h"

using namespace refgeom;

namespace {

enum { OK = 0, E_NUMBER = 1, E_FLAG = 2, E_COMMAND = 3, E_MISSING_MOVE_TO = 4, E_NON_ASCII = 5 };

struct Source {
    const uint8_t* s;
    uint32_t n, pos = 0;
    uint32_t current = ' ';
    int line = 0, col = 0;
    bool finished = false;
    bool non_ascii = false;
    int na_line = 0, na_col = 0;

    void take(uint32_t c) {
        if (c >= 0x80 && !non_ascii) {
            non_ascii = true;
            na_line = line;
            na_col = col;
        }
        current = c;
    }
    Source(const uint8_t* text, uint32_t len) : s(text), n(len) {
        // Source::with_position(0, 0, ..) (parser.rs:66-86)
        if (n > 0) {
            take(s[0]);
            pos = 1;
            finished = false;
        } else {
            current = ' ';
            finished = true;
        }
        line = current == '\n' ? 1 : 0;
        col = 0;
    }
    // parser.rs:100-124
    void advance_one() {
        if (finished) return;
        if (pos < n) {
            uint32_t c = s[pos++];
            if (c == '\n') {
                current = '\n';
                line += 1;
                col = -1;
            } else {
                col += 1;
                take(c);
            }
        } else {
            current = '~';
            finished = true;
        }
    }
    static bool is_ws(uint32_t c) { return c == ' ' || (c >= 9 && c <= 13); }
    void skip_whitespace() {
        while (!finished && (is_ws(current) || current == ',')) advance_one();
    }
};

bool digit(uint32_t c) { return c >= '0' && c <= '9'; }

// Rust's f32 grammar over the collected text (no '+' can be collected).
bool rust_f32(const std::string& t, float* out) {
    size_t i = 0;
    if (i < t.size() && t[i] == '-') ++i;
    size_t m = 0;
    while (i < t.size() && digit((uint8_t)t[i])) ++i, ++m;
    if (i < t.size() && t[i] == '.') {
        ++i;
        while (i < t.size() && digit((uint8_t)t[i])) ++i, ++m;
    }
    if (m == 0) return false;
    if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
        ++i;
        if (i < t.size() && (t[i] == '-' || t[i] == '+')) ++i;
        size_t e = 0;
        while (i < t.size() && digit((uint8_t)t[i])) ++i, ++e;
        if (e == 0) return false;
    }
    if (i != t.size()) return false;
    *out = std::strtof(t.c_str(), nullptr);
    return true;
}

struct Builder {  // BuilderImpl (path.rs:479-545)
    std::vector<V2> points;
    std::vector<uint32_t> verbs;
    V2 first{0.0f, 0.0f};
    void begin(V2 p) {
        first = p;
        points.push_back(p);
        verbs.push_back(3);
    }
    void end(bool close) {
        if (close) points.push_back(first);
        verbs.push_back(close ? 4 : 5);
    }
    void line_to(V2 p) {
        points.push_back(p);
        verbs.push_back(0);
    }
    void quad_to(V2 c, V2 p) {
        points.push_back(c);
        points.push_back(p);
        verbs.push_back(1);
    }
    void cubic_to(V2 c1, V2 c2, V2 p) {
        points.push_back(c1);
        points.push_back(c2);
        points.push_back(p);
        verbs.push_back(2);
    }
};

struct Err {
    int kind = OK, line = 0, col = 0;
    uint32_t ch = 0;
};

struct Parser {
    Source& src;
    Builder& out;
    Err& err;
    int num_attributes;
    V2 current_position{0.0f, 0.0f};

    // parse_number (parser.rs:421-466)
    bool number(float* v) {
        std::string buf;
        src.skip_whitespace();
        const int line = src.line, col = src.col;
        if (src.current == '-') {
            buf.push_back('-');
            src.advance_one();
        }
        while (digit(src.current)) {
            buf.push_back((char)src.current);
            src.advance_one();
        }
        if (src.current == '.') {
            buf.push_back('.');
            src.advance_one();
            while (digit(src.current)) {
                buf.push_back((char)src.current);
                src.advance_one();
            }
        }
        if (src.current == 'e' || src.current == 'E') {
            buf.push_back((char)src.current);
            src.advance_one();
            if (src.current == '-') {
                buf.push_back('-');
                src.advance_one();
            }
            while (digit(src.current)) {
                buf.push_back((char)src.current);
                src.advance_one();
            }
        }
        if (rust_f32(buf, v)) return true;
        err.kind = E_NUMBER;
        err.line = line;
        err.col = col;
        return false;
    }
    // parse_flag (parser.rs:468-489)
    bool flag(bool* f) {
        src.skip_whitespace();
        if (src.current == '1' || src.current == '0') {
            *f = src.current == '1';
            src.advance_one();
            return true;
        }
        err.kind = E_FLAG;
        err.line = src.line;
        err.col = src.col;
        err.ch = src.current;
        return false;
    }
    bool attributes() {
        float d;
        for (int k = 0; k < num_attributes; ++k)
            if (!number(&d)) return false;
        return true;
    }
    // parse_point (parser.rs:403-419)
    bool point(bool rel, V2* p) {
        float x, y;
        if (!number(&x) || !number(&y)) return false;
        if (rel) {
            x += current_position.x;
            y += current_position.y;
        }
        *p = v2(x, y);
        return true;
    }
    // parse_endpoint (parser.rs:376-387)
    bool endpoint(bool rel, V2* p) {
        if (!point(rel, p)) return false;
        current_position = *p;
        return attributes();
    }
};

float cr_cos(float x) { return (float)std::cos((double)x); }
float cr_sin(float x) { return (float)std::sin((double)x); }
float cr_tan(float x) { return (float)std::tan((double)x); }

// euclid Rotation2D::transform_vector
V2 cr_rotate(float angle, V2 p) {
    const float s = cr_sin(angle), c = cr_cos(angle);
    return v2(p.x * c - p.y * s, p.y * c + p.x * s);
}

// arc_to_quadratic_beziers_with_t (arc.rs:668-701) with sample_ellipse (:741-746) and
// tangent_at_angle (:510-516): f(ctrl, to) per quadratic.
template <class F>
void cr_arc_to_quads(const Arc& arc, F&& f) {
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_4 = 0.785398163397448309616f;
    float sign = std::copysign(1.0f, arc.sweep_angle);
    if (std::isnan(arc.sweep_angle)) sign = arc.sweep_angle;
    const float sweep = fminr(std::fabs(arc.sweep_angle), PI * 2.0f);
    const float n_steps = std::ceil(sweep / FRAC_PI_4);
    const float step = sweep / n_steps * sign;
    const float alpha = cr_tan(step * 0.5f);
    const int n = (int)n_steps;
    for (int i = 0; i < n; ++i) {
        const float a1 = arc.start_angle + step * (float)i;
        const float a2 = arc.start_angle + step * (float)(i + 1);
        const V2 from = arc.center + cr_rotate(arc.x_rotation, v2(arc.radii.x * cr_cos(a1), arc.radii.y * cr_sin(a1)));
        const V2 to = arc.center + cr_rotate(arc.x_rotation, v2(arc.radii.x * cr_cos(a2), arc.radii.y * cr_sin(a2)));
        const V2 tangent = cr_rotate(arc.x_rotation, v2(-arc.radii.x * cr_sin(a1), arc.radii.y * cr_cos(a1)));
        f(from + tangent * alpha, to);
    }
}

// Arc::from_svg_arc (arc.rs:58-146)
Arc svg_to_arc(V2 from, V2 to, V2 radii, float x_rotation, bool large_arc, bool sweep) {
    const float PI = 3.14159265358979323846f;
    float rx = std::fabs(radii.x), ry = std::fabs(radii.y);
    const float xr = std::fmod(x_rotation, 2.0f * PI);
    const float cos_phi = cr_cos(xr), sin_phi = cr_sin(xr);
    const float hd_x = (from.x - to.x) / 2.0f, hd_y = (from.y - to.y) / 2.0f;
    const float hs_x = (from.x + to.x) / 2.0f, hs_y = (from.y + to.y) / 2.0f;
    const V2 p = v2(cos_phi * hd_x + sin_phi * hd_y, -sin_phi * hd_x + cos_phi * hd_y);
    const float rf = p.x * p.x / (rx * rx) + p.y * p.y / (ry * ry);
    if (rf > 1.0f) {
        const float scale = std::sqrt(rf);
        rx *= scale;
        ry *= scale;
    }
    const float rxry = rx * ry, rxpy = rx * p.y, rypx = ry * p.x;
    const float sum_of_sq = rxpy * rxpy + rypx * rypx;
    const float sign_coe = large_arc == sweep ? -1.0f : 1.0f;
    const float coe = sign_coe * std::sqrt(std::fabs((rxry * rxry - sum_of_sq) / sum_of_sq));
    const float tcx = coe * rxpy / ry;
    const float tcy = -coe * rypx / rx;
    const V2 center = v2(cos_phi * tcx - sin_phi * tcy + hs_x, sin_phi * tcx + cos_phi * tcy + hs_y);
    const V2 sv = v2((p.x - tcx) / rx, (p.y - tcy) / ry);
    const V2 ev = v2((-p.x - tcx) / rx, (-p.y - tcy) / ry);
    const float two_pi = 2.0f * PI;
    const float start_angle = angle_from_x_axis(sv);
    float sweep_angle = std::fmod(angle_from_x_axis(ev) - start_angle, two_pi);
    if (sweep && sweep_angle < 0.0f) sweep_angle += two_pi;
    else if (!sweep && sweep_angle > 0.0f) sweep_angle -= two_pi;
    Arc a;
    a.center = center;
    a.radii = v2(rx, ry);
    a.start_angle = start_angle;
    a.sweep_angle = sweep_angle;
    a.x_rotation = x_rotation;
    return a;
}

// PathParser::parse (parser.rs:154-175) + parse_path (:177-357)
void parse(const uint8_t* text, uint32_t len, int num_attributes, int stop_at, Builder& out,
           Err& err) {
    Source src(text, len);
    Parser P{src, out, err, num_attributes};
    bool need_end = false;
    V2 first_position{0.0f, 0.0f};
    bool need_start = false;
    bool has_pc = false, has_pq = false;
    V2 pc{0, 0}, pq{0, 0};
    uint32_t implicit_cmd = 'M';
    src.skip_whitespace();
    while (!src.finished) {
        uint32_t cmd = src.current;
        const int cmd_line = src.line, cmd_col = src.col;
        if (stop_at >= 0 && (uint32_t)stop_at == cmd) break;
        const bool alpha = (cmd >= 'a' && cmd <= 'z') || (cmd >= 'A' && cmd <= 'Z');
        if (alpha) src.advance_one();
        else cmd = implicit_cmd;
        if (need_start && cmd != 'm' && cmd != 'M') {
            err.kind = E_MISSING_MOVE_TO;
            err.line = cmd_line;
            err.col = cmd_col;
            err.ch = cmd;
            break;
        }
        const bool rel = cmd >= 'a' && cmd <= 'z';
        V2 to, c1, c2;
        bool ok = true;
        switch (cmd) {
            case 'l':
            case 'L':
                ok = P.endpoint(rel, &to);
                if (ok) out.line_to(to);
                break;
            case 'h':
            case 'H': {
                float x;
                ok = P.number(&x);
                if (!ok) break;
                if (rel) x += P.current_position.x;
                to = v2(x, P.current_position.y);
                P.current_position = to;
                ok = P.attributes();
                if (ok) out.line_to(to);
                break;
            }
            case 'v':
            case 'V': {
                float y;
                ok = P.number(&y);
                if (!ok) break;
                if (rel) y += P.current_position.y;
                to = v2(P.current_position.x, y);
                P.current_position = to;
                ok = P.attributes();
                if (ok) out.line_to(to);
                break;
            }
            case 'q':
            case 'Q':
                ok = P.point(rel, &c1) && P.endpoint(rel, &to);
                if (!ok) break;
                pq = c1;
                has_pq = true;
                out.quad_to(c1, to);
                break;
            case 't':
            case 'T': {
                // get_smooth_ctrl (parser.rs:359-365)
                const V2 cp = P.current_position;
                c1 = has_pq ? cp + (cp - pq) : cp;
                ok = P.endpoint(rel, &to);
                if (!ok) break;
                pq = c1;
                has_pq = true;
                out.quad_to(c1, to);
                break;
            }
            case 'c':
            case 'C':
                ok = P.point(rel, &c1) && P.point(rel, &c2) && P.endpoint(rel, &to);
                if (!ok) break;
                pc = c2;
                has_pc = true;
                out.cubic_to(c1, c2, to);
                break;
            case 's':
            case 'S': {
                const V2 cp = P.current_position;
                c1 = has_pc ? cp + (cp - pc) : cp;
                ok = P.point(rel, &c2) && P.endpoint(rel, &to);
                if (!ok) break;
                pc = c2;
                has_pc = true;
                out.cubic_to(c1, c2, to);
                break;
            }
            case 'a':
            case 'A': {
                const V2 from = P.current_position;
                float rx, ry, xdeg;
                bool large_arc, sweep;
                ok = P.number(&rx) && P.number(&ry) && P.number(&xdeg) && P.flag(&large_arc) &&
                     P.flag(&sweep) && P.endpoint(rel, &to);
                if (!ok) break;
                // Angle::degrees -> f32::to_radians: x * (PI / 180)
                const float xrot = xdeg * (3.14159265358979323846f / 180.0f);
                const float EPS = 1.1920929e-07f;
                if (std::fabs(rx) <= EPS || std::fabs(ry) <= EPS || from == to) {
                    out.line_to(to);
                } else {
                    const Arc arc = svg_to_arc(from, to, v2(rx, ry), xrot, large_arc, sweep);
                    cr_arc_to_quads(arc, [&](V2 c, V2 t) { out.quad_to(c, t); });
                }
                break;
            }
            case 'm':
            case 'M':
                if (need_end) {
                    out.end(false);
                    need_end = false;
                }
                ok = P.endpoint(rel, &to);
                if (!ok) break;
                first_position = to;
                out.begin(to);
                need_end = true;
                need_start = false;
                break;
            case 'z':
            case 'Z':
                out.end(true);
                P.current_position = first_position;
                need_end = false;
                need_start = true;
                break;
            default:
                err.kind = E_COMMAND;
                err.line = cmd_line;
                err.col = cmd_col;
                err.ch = cmd;
                ok = false;
        }
        if (!ok) break;
        if (cmd == 'c' || cmd == 'C' || cmd == 's' || cmd == 'S') {
            has_pq = false;
        } else if (cmd == 'q' || cmd == 'Q' || cmd == 't' || cmd == 'T') {
            has_pc = false;
        } else {
            has_pc = has_pq = false;
        }
        implicit_cmd = cmd == 'm' ? 'l' : cmd == 'M' ? 'L' : cmd == 'z' ? 'm' : cmd == 'Z' ? 'M' : cmd;
        src.skip_whitespace();
    }
    if (need_end) out.end(false);
    if (src.non_ascii) {
        err.kind = E_NON_ASCII;
        err.line = src.na_line;
        err.col = src.na_col;
        err.ch = 0;
    }
}

}  // namespace

extern "C" {

// Parses one string; writes up to the capacities and returns the counts in n_verbs /
// n_points (call again with larger arrays when they exceed the capacities).
void orc_parse_svg(const uint8_t* text, uint32_t len, int num_attributes, int stop_at,
                   uint8_t* verbs, uint32_t verb_cap, float* points, uint32_t point_cap,
                   uint32_t* n_verbs, uint32_t* n_points, int32_t* error) {
    Builder b;
    Err e;
    parse(text, len, num_attributes, stop_at, b, e);
    *n_verbs = (uint32_t)b.verbs.size();
    *n_points = (uint32_t)b.points.size();
    for (size_t i = 0; i < b.verbs.size() && i < verb_cap; ++i) verbs[i] = (uint8_t)b.verbs[i];
    for (size_t i = 0; i < b.points.size() && i < point_cap; ++i) {
        points[2 * i] = b.points[i].x;
        points[2 * i + 1] = b.points[i].y;
    }
    error[0] = e.kind;
    error[1] = e.line;
    error[2] = e.col;
    error[3] = (int32_t)e.ch;
}

// str::parse::<f32> of a string (the grammar check and strtof above); returns 1 when valid.
int orc_parse_f32(const char* s, float* out) { return rust_f32(std::string(s), out) ? 1 : 0; }

}  // extern "C"
