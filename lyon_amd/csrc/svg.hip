// SVG path data on the device: lyon_extra's PathParser (crates/extra/src/parser.rs:140-462)
// writing into lyon_path's Path layout (BuilderImpl, crates/path/src/path.rs:516-611), one
// lane per path string.  The parse is a serial state machine over the characters, so a
// lane walks one string, writing into a staging room sized by the string's length; scans
// of the counts place the paths end to end and a copy kernel moves them there.
//
// Numbers: the parser collects `-? digits* (. digits*)? ([eE] -? digits*)?` and converts it
// with Rust's `str::parse::<f32>`, which is correctly rounded (round half to even) for
// every input.  Here: up to 19 significant digits are kept in a u64 (more set a sticky
// bit); when they fit 2^53 and the decimal exponent is within ±22, one IEEE double
// operation gives the value correctly rounded to double, and its rounding to f32 is the
// correct one unless the double lies exactly on an f32 midpoint; every other case is
// rounded exactly with 320-bit integer arithmetic, longer digit strings by comparing their
// remaining digits with the decimal expansion of the midpoint they straddle.
//
// Arcs: SvgArc::to_arc (Arc::from_svg_arc, crates/geom/src/arc.rs:58-146) and
// Arc::for_each_quadratic_bezier_with_t (arc.rs:668-701), as the parser does
// (parser.rs:287-315).  Only ASCII input is accepted: a non-ASCII byte reached by the
// parser ends the path with the NonAscii error.
#include <hip/hip_runtime.h>

#include "arc_core.h"
#include "lyg_math.h"
#include "lyg_types.h"

namespace lyg {

namespace {

// ---- decimal to f32 -------------------------------------------------------------------
constexpr int BW = 10;  // 320-bit unsigned integers, little-endian 32-bit words
struct Big {
    uint32_t w[BW];
};

__device__ void big_set(Big& b, uint64_t v) {
#pragma unroll
    for (int i = 0; i < BW; ++i) b.w[i] = 0;
    b.w[0] = (uint32_t)v;
    b.w[1] = (uint32_t)(v >> 32);
}
__device__ void big_mul(Big& b, uint32_t m) {
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < BW; ++i) {
        const uint64_t t = (uint64_t)b.w[i] * m + c;
        b.w[i] = (uint32_t)t;
        c = t >> 32;
    }
}
__device__ void big_pow10(Big& b, int e) {
    const uint32_t P[9] = {1u, 10u, 100u, 1000u, 10000u, 100000u, 1000000u, 10000000u,
                           100000000u};
    while (e >= 9) {
        big_mul(b, 1000000000u);
        e -= 9;
    }
    if (e > 0) big_mul(b, P[e]);
}
__device__ void big_shl(Big& b, int s) {
    const int ws = s >> 5, bs = s & 31;
    for (int i = BW - 1; i >= 0; --i) {
        const int j = i - ws;
        const uint32_t hi = j >= 0 ? b.w[j] : 0u;
        const uint32_t lo = j - 1 >= 0 ? b.w[j - 1] : 0u;
        b.w[i] = bs ? (hi << bs) | (lo >> (32 - bs)) : hi;
    }
}
__device__ int big_cmp(const Big& a, const Big& b) {
    for (int i = BW - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}
__device__ void big_sub(Big& a, const Big& b) {
    int64_t br = 0;
#pragma unroll
    for (int i = 0; i < BW; ++i) {
        int64_t t = (int64_t)a.w[i] - (int64_t)b.w[i] - br;
        br = t < 0 ? 1 : 0;
        a.w[i] = (uint32_t)(t + (br << 32));
    }
}
__device__ int big_bits(const Big& a) {
    for (int i = BW - 1; i >= 0; --i)
        if (a.w[i]) return 32 * i + 32 - __builtin_clz(a.w[i]);
    return 0;
}

// f32 bits of w * 10^e correctly rounded; sticky +1 / -1 says the exact value lies just
// above / below w * 10^e (it only breaks ties).  Needs w > 0 and the value within
// [10^-46, 10^39].
__device__ uint32_t exact_round(uint64_t w, int e, int sticky) {
    Big N, D;
    big_set(N, w);
    big_set(D, 1);
    if (e >= 0) big_pow10(N, e);
    else big_pow10(D, -e);
    int k = big_bits(N) - big_bits(D);  // N / D in (2^(k-1), 2^(k+1))
    {
        Big a = N, b = D;
        if (k >= 0) big_shl(b, k);
        else big_shl(a, -k);
        if (big_cmp(a, b) < 0) --k;  // now 2^k <= N / D < 2^(k+1)
    }
    if (k > 127) return 0x7f800000u;
    const int s = k >= -126 ? 23 - k : 149;
    if (s >= 0) big_shl(N, s);
    else big_shl(D, -s);
    uint32_t q = 0;  // floor(N / D) < 2^24
    for (int bit = 24; bit >= 0; --bit) {
        Big t = D;
        big_shl(t, bit);
        if (big_cmp(N, t) >= 0) {
            big_sub(N, t);
            q |= 1u << bit;
        }
    }
    big_shl(N, 1);  // 2 * remainder against D
    int c = big_cmp(N, D);
    if (c == 0) c = sticky != 0 ? sticky : ((q & 1u) ? 1 : -1);
    if (c > 0) ++q;
    if (k >= -126) {
        if (q == (1u << 24)) {
            q = 1u << 23;
            ++k;
        }
        if (k > 127) return 0x7f800000u;
        return ((uint32_t)(k + 127) << 23) | (q - (1u << 23));
    }
    return q;  // subnormal; q == 2^23 is the smallest normal number
}

// w * 10^e through one double operation (w <= 2^53, |e| <= 22: both operands exact, so the
// double is correctly rounded); false when the double is an f32 midpoint.
__device__ bool fast_round(uint64_t w, int e, uint32_t* bits) {
    if (w > (1ull << 53) || e < -22 || e > 22) return false;
    double p = 1.0;
    for (int i = 0; i < (e < 0 ? -e : e); ++i) p *= 10.0;  // exact up to 10^22
    const double d = e >= 0 ? (double)w * p : (double)w / p;
    const float lo = __double2float_rz(d);
    const double dl = (double)lo;
    if (dl == d) {
        *bits = __float_as_uint(lo);
        return true;
    }
    const uint32_t hb = __float_as_uint(lo) + 1u;
    const double mid = (dl + (double)__uint_as_float(hb)) * 0.5;
    if (d == mid) return false;
    *bits = d < mid ? __float_as_uint(lo) : hb;
    return true;
}

__device__ __forceinline__ bool is_digit(uint32_t c) { return c >= '0' && c <= '9'; }

// Beyond 19 significant digits, when w * 10^e and (w + 1) * 10^e round to different f32
// values lo and hi (then adjacent: the interval is 1e-18 of the value wide), the value
// rounds by where it lies against their midpoint m = mm * 2^q.  With r = m / 10^e - w in
// (0, 1) as the fraction rn / rd, the remaining digits d20 d21 ... (the fraction 0.d20d21...
// of the value over 10^e) are compared with r's decimal digits one at a time; r is dyadic
// over a power of ten, so its expansion ends.  Returns lo or hi (ties to even).
__device__ uint32_t decide_midpoint(const uint8_t* s, uint32_t a, uint32_t b, uint64_t w, int e,
                                    uint32_t lo) {
    uint64_t mm;
    int q;
    if ((lo >> 23) == 0) {  // zero or subnormal: lo * 2^-149, the midpoint (2 lo + 1) 2^-150
        mm = 2ull * lo + 1;
        q = -150;
    } else {
        mm = 2ull * ((lo & 0x7fffffu) | 0x800000u) + 1;
        q = (int)(lo >> 23) - 151;
    }
    // r = mm * 2^q * 10^-e - w = rn / rd, rd = 2^max(-q, 0) * 10^max(e, 0)
    Big rn, rd, t;
    big_set(rn, mm);
    if (q > 0) big_shl(rn, q);
    if (e < 0) big_pow10(rn, -e);
    big_set(rd, 1);
    if (q < 0) big_shl(rd, -q);
    if (e > 0) big_pow10(rd, e);
    t = rd;
    big_mul(t, (uint32_t)(w >> 32));
    big_shl(t, 32);
    {
        Big u = rd;
        big_mul(u, (uint32_t)w);
        // t += u
        uint64_t c = 0;
        for (int i = 0; i < BW; ++i) {
            const uint64_t v = (uint64_t)t.w[i] + u.w[i] + c;
            t.w[i] = (uint32_t)v;
            c = v >> 32;
        }
    }
    big_sub(rn, t);  // w * rd <= m * rd / 10^e: no borrow
    // the digits after the 19th significant one, in order
    int sig = 0;
    for (uint32_t i = a; i < b; ++i) {
        const uint32_t ch = s[i];
        if (ch == 'e' || ch == 'E') break;
        if (!is_digit(ch)) continue;
        const uint32_t d = ch - '0';
        if (sig == 0 && d == 0) continue;  // leading zeros
        if (++sig <= 19) continue;
        big_mul(rn, 10);  // next digit of r: floor(10 rn / rd), at most 9
        uint32_t dm = 0;
        while (big_cmp(rn, rd) >= 0) {
            big_sub(rn, rd);
            ++dm;
        }
        if (d != dm) return d > dm ? lo + 1 : lo;
    }
    bool zero = true;
    for (int i = 0; i < BW; ++i) zero = zero && rn.w[i] == 0;
    if (!zero) return lo;          // the digits end below m
    return (lo & 1u) ? lo + 1 : lo;  // exactly m: to even
}

enum { NUM_OK = 0, NUM_SYNTAX = 1 };

// str::parse::<f32> of s[a, b), which holds what the parser collected.
__device__ int parse_f32(const uint8_t* s, uint32_t a, uint32_t b, float* out) {
    uint32_t i = a;
    bool neg = false;
    if (i < b && s[i] == '-') {
        neg = true;
        ++i;
    }
    uint64_t w = 0;
    int kept = 0, mant = 0;
    int64_t shift = 0;  // decimal exponent contributed by the digits
    bool sticky = false;
    while (i < b && is_digit(s[i])) {
        const uint32_t d = s[i] - '0';
        ++mant;
        if (kept == 0 && d == 0) {
        } else if (kept < 19) {
            w = w * 10 + d;
            ++kept;
        } else {
            ++shift;
            sticky |= d != 0;
        }
        ++i;
    }
    if (i < b && s[i] == '.') {
        ++i;
        while (i < b && is_digit(s[i])) {
            const uint32_t d = s[i] - '0';
            ++mant;
            if (kept == 0 && d == 0) {
                --shift;
            } else if (kept < 19) {
                w = w * 10 + d;
                ++kept;
                --shift;
            } else {
                sticky |= d != 0;
            }
            ++i;
        }
    }
    if (mant == 0) return NUM_SYNTAX;
    int64_t ex = 0;
    if (i < b && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        bool eneg = false;
        if (i < b && s[i] == '-') {
            eneg = true;
            ++i;
        }
        int ed = 0;
        while (i < b && is_digit(s[i])) {
            ex = ex * 10 + (s[i] - '0');
            if (ex > 1000000) ex = 1000000;
            ++ed;
            ++i;
        }
        if (ed == 0) return NUM_SYNTAX;
        if (eneg) ex = -ex;
    }
    if (i != b) return NUM_SYNTAX;
    uint32_t bits;
    const int64_t e64 = ex + shift;
    if (w == 0 || e64 + kept <= -46) {
        bits = 0;
    } else if (e64 + kept > 39) {
        bits = 0x7f800000u;
    } else {
        const int e = (int)e64;
        if (sticky) {
            const uint32_t lo = exact_round(w, e, 1);
            const uint32_t hi = exact_round(w + 1, e, -1);
            bits = lo == hi ? lo : decide_midpoint(s, a, b, w, e, lo);
        } else if (!fast_round(w, e, &bits)) {
            bits = exact_round(w, e, 0);
        }
    }
    *out = __uint_as_float(bits | (neg ? 0x80000000u : 0u));
    return NUM_OK;
}

// ---- the character source (parser.rs:48-126) ------------------------------------------
struct Src {
    const uint8_t* s;
    uint32_t pos, end;  // next byte to read, end of the string
    uint32_t at;        // byte index of `cur` (end once finished)
    uint32_t cur;
    int line, col;
    bool finished;
    bool non_ascii;
    int na_line, na_col;
};

__device__ void src_take(Src& S, uint32_t c) {
    if (c >= 0x80u && !S.non_ascii) {
        S.non_ascii = true;
        S.na_line = S.line;
        S.na_col = S.col;
    }
    S.cur = c;
}
__device__ void src_init(Src& S, const uint8_t* s, uint32_t a, uint32_t b) {
    S.s = s;
    S.end = b;
    S.non_ascii = false;
    S.na_line = S.na_col = 0;
    S.line = 0;
    S.col = 0;
    if (a < b) {
        S.at = a;
        S.pos = a + 1;
        S.finished = false;
        src_take(S, s[a]);
        if (S.cur == '\n') S.line = 1;
    } else {
        S.at = b;
        S.pos = b;
        S.cur = ' ';
        S.finished = true;
    }
}
__device__ void src_advance(Src& S) {
    if (S.finished) return;
    if (S.pos < S.end) {
        const uint32_t c = S.s[S.pos];
        S.at = S.pos;
        ++S.pos;
        if (c == '\n') {
            S.cur = '\n';
            ++S.line;
            S.col = -1;
        } else {
            ++S.col;
            src_take(S, c);
        }
    } else {
        S.cur = '~';
        S.finished = true;
        S.at = S.end;
    }
}
__device__ __forceinline__ bool is_ws(uint32_t c) { return c == ' ' || (c >= 9 && c <= 13); }
__device__ void skip_ws(Src& S) {
    while (!S.finished && (is_ws(S.cur) || S.cur == ',')) src_advance(S);
}

// ---- the output: BuilderImpl (path.rs:516-611) ----------------------------------------
template <bool EMIT>
struct Out {
    uint8_t* verbs;
    float2* points;
    uint32_t nv, np;
    vec2 first;
    uint32_t cap_v, cap_p;  // staging room of the string (writes beyond it are dropped)
    __device__ void pt(vec2 p) {
        if (EMIT && np < cap_p) points[np] = make_float2(p.x, p.y);
        ++np;
    }
    __device__ void verb(uint32_t v) {
        if (EMIT && nv < cap_v) verbs[nv] = (uint8_t)v;
        ++nv;
    }
    __device__ void begin(vec2 p) {
        first = p;
        pt(p);
        verb(3);
    }
    __device__ void end(bool close) {
        if (close) pt(first);
        verb(close ? 4 : 5);
    }
    __device__ void line_to(vec2 p) {
        pt(p);
        verb(0);
    }
    __device__ void quad_to(vec2 c, vec2 p) {
        pt(c);
        pt(p);
        verb(1);
    }
    __device__ void cubic_to(vec2 c1, vec2 c2, vec2 p) {
        pt(c1);
        pt(c2);
        pt(p);
        verb(2);
    }
};

struct PErr {
    int32_t kind, line, col;
    uint32_t ch;
};

__device__ bool parse_number(Src& S, float* v, PErr& E) {
    skip_ws(S);
    const int line = S.line, col = S.col;
    const uint32_t a = S.at;
    if (S.cur == '-') src_advance(S);
    while (is_digit(S.cur)) src_advance(S);
    if (S.cur == '.') {
        src_advance(S);
        while (is_digit(S.cur)) src_advance(S);
    }
    if (S.cur == 'e' || S.cur == 'E') {
        src_advance(S);
        if (S.cur == '-') src_advance(S);
        while (is_digit(S.cur)) src_advance(S);
    }
    const int r = parse_f32(S.s, a, S.at, v);
    if (r == NUM_OK) return true;
    E = PErr{LYG_SVG_ERR_NUMBER, line, col, 0u};
    return false;
}

__device__ bool parse_flag(Src& S, bool* f, PErr& E) {
    skip_ws(S);
    if (S.cur == '1' || S.cur == '0') {
        *f = S.cur == '1';
        src_advance(S);
        return true;
    }
    E = PErr{LYG_SVG_ERR_FLAG, S.line, S.col, S.cur};
    return false;
}

// SvgArc::to_arc (Arc::from_svg_arc, arc.rs:58-146)
__device__ ArcD svg_arc_to_arc(vec2 from, vec2 to, vec2 radii, float xrot, bool large_arc,
                               bool sweep) {
    const float PI = 3.14159265358979323846f;
    const float two_pi = 2.0f * PI;
    float rx = ffabs(radii.x), ry = ffabs(radii.y);
    const float xr = fmodf(xrot, two_pi);
    const float cos_phi = t_cos(xr), sin_phi = t_sin(xr);
    const float hd_x = (from.x - to.x) / 2.0f, hd_y = (from.y - to.y) / 2.0f;
    const float hs_x = (from.x + to.x) / 2.0f, hs_y = (from.y + to.y) / 2.0f;
    const vec2 p = mk(cos_phi * hd_x + sin_phi * hd_y, -sin_phi * hd_x + cos_phi * hd_y);
    const float rf = p.x * p.x / (rx * rx) + p.y * p.y / (ry * ry);
    if (rf > 1.0f) {
        const float scale = fsqrt(rf);
        rx *= scale;
        ry *= scale;
    }
    const float rxry = rx * ry, rxpy = rx * p.y, rypx = ry * p.x;
    const float sum_of_sq = rxpy * rxpy + rypx * rypx;
    const float sign_coe = large_arc == sweep ? -1.0f : 1.0f;
    const float coe = sign_coe * fsqrt(ffabs((rxry * rxry - sum_of_sq) / sum_of_sq));
    const float tcx = coe * rxpy / ry;
    const float tcy = -coe * rypx / rx;
    const vec2 center = mk(cos_phi * tcx - sin_phi * tcy + hs_x, sin_phi * tcx + cos_phi * tcy + hs_y);
    const vec2 start_v = mk((p.x - tcx) / rx, (p.y - tcy) / ry);
    const vec2 end_v = mk((-p.x - tcx) / rx, (-p.y - tcy) / ry);
    const float start_angle = angle_of(start_v);
    float sweep_angle = fmodf(angle_of(end_v) - start_angle, two_pi);
    if (sweep && sweep_angle < 0.0f) sweep_angle += two_pi;
    else if (!sweep && sweep_angle > 0.0f) sweep_angle -= two_pi;
    return ArcD{center, mk(rx, ry), start_angle, sweep_angle, xrot};
}

// Arguments a command reads (parser.rs:225-337): L/M/T 2, H/V 1, Q/S 4, C 6, A 7 (flags at
// 3 and 4), Z 0; -1 for a character that is no command.  Endpoint attributes come on top.
__device__ __forceinline__ int cmd_args(uint32_t c) {
    switch (c | 0x20u) {
        case 'm':
        case 'l':
        case 't':
            return 2;
        case 'h':
        case 'v':
            return 1;
        case 'q':
        case 's':
            return 4;
        case 'c':
            return 6;
        case 'a':
            return 7;
        case 'z':
            return 0;
        default:
            return -1;
    }
}

// PathParser::parse (parser.rs:154-175) and parse_path (:177-357), driven by tokens: every
// iteration of the loop reads one number or flag, so the lanes of a wave (one string each)
// spend their time in the same code whatever command each of them is in; a command is
// applied when its last argument has been read.  Relative coordinates, smooth control
// points and an arc's start use the position at the command's start, as parse_point /
// get_smooth_ctrl do (the position only changes at the command's endpoint).
template <bool EMIT>
__device__ void parse_svg(const uint8_t* s, uint32_t a, uint32_t b, int nattr, int stop_at,
                          Out<EMIT>& O, PErr& E) {
    Src S;
    src_init(S, s, a, b);
    E = PErr{LYG_SVG_OK, 0, 0, 0u};
    bool need_end = false;
    vec2 cur = mk(0.0f, 0.0f), first_pos = mk(0.0f, 0.0f);
    bool need_start = false;
    bool has_pc = false, has_pq = false;
    vec2 pc = cur, pq = cur;
    uint32_t implicit = 'M';
    uint32_t cmd = 0;
    int k = 0, nargs = 0, nums = 0;  // next argument, arguments in all, of them before attrs
    float g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0, g5 = 0, g6 = 0;
    skip_ws(S);
    for (;;) {
        if (k == 0) {
            // a command boundary
            if (S.finished) break;
            uint32_t c = S.cur;
            const int cl = S.line, cc = S.col;
            if (stop_at >= 0 && (uint32_t)stop_at == c) break;
            if ((c | 0x20u) >= 'a' && (c | 0x20u) <= 'z') src_advance(S);
            else c = implicit;
            if (need_start && c != 'm' && c != 'M') {
                E = PErr{LYG_SVG_ERR_MISSING_MOVE_TO, cl, cc, c};
                break;
            }
            const int na = cmd_args(c);
            if (na < 0) {
                E = PErr{LYG_SVG_ERR_COMMAND, cl, cc, c};
                break;
            }
            cmd = c;
            if (na == 0) {
                // Z
                O.end(true);
                cur = first_pos;
                need_end = false;
                need_start = true;
                has_pc = has_pq = false;
                implicit = cmd == 'z' ? 'm' : 'M';
                skip_ws(S);
                continue;
            }
            if ((cmd | 0x20u) == 'm' && need_end) {
                O.end(false);
                need_end = false;
            }
            nums = na;
            nargs = na + nattr;
        }
        // one argument: a flag (an arc's 4th and 5th) or a number
        float v = 0.0f;
        if ((cmd | 0x20u) == 'a' && (k == 3 || k == 4)) {
            bool f;
            if (!parse_flag(S, &f, E)) break;
            v = f ? 1.0f : 0.0f;
        } else {
            if (!parse_number(S, &v, E)) break;
        }
        if (k < nums) {
            g0 = k == 0 ? v : g0;
            g1 = k == 1 ? v : g1;
            g2 = k == 2 ? v : g2;
            g3 = k == 3 ? v : g3;
            g4 = k == 4 ? v : g4;
            g5 = k == 5 ? v : g5;
            g6 = k == 6 ? v : g6;
        }
        ++k;
        if (k < nargs) continue;
        // the command's arguments are all read: apply it
        k = 0;
        const bool rel = cmd >= 'a' && cmd <= 'z';
        const vec2 o = rel ? cur : mk(0.0f, 0.0f);
        auto P = [&](float x, float y) { return rel ? mk(x + o.x, y + o.y) : mk(x, y); };
        switch (cmd | 0x20u) {
            case 'l': {
                const vec2 to = P(g0, g1);
                cur = to;
                O.line_to(to);
                break;
            }
            case 'h': {
                const vec2 to = mk(rel ? g0 + cur.x : g0, cur.y);
                cur = to;
                O.line_to(to);
                break;
            }
            case 'v': {
                const vec2 to = mk(cur.x, rel ? g0 + cur.y : g0);
                cur = to;
                O.line_to(to);
                break;
            }
            case 'q': {
                const vec2 c1 = P(g0, g1), to = P(g2, g3);
                cur = to;
                pq = c1;
                has_pq = true;
                O.quad_to(c1, to);
                break;
            }
            case 't': {
                const vec2 c1 = has_pq ? add(cur, sub(cur, pq)) : cur;
                const vec2 to = P(g0, g1);
                cur = to;
                pq = c1;
                has_pq = true;
                O.quad_to(c1, to);
                break;
            }
            case 'c': {
                const vec2 c1 = P(g0, g1), c2 = P(g2, g3), to = P(g4, g5);
                cur = to;
                pc = c2;
                has_pc = true;
                O.cubic_to(c1, c2, to);
                break;
            }
            case 's': {
                const vec2 c1 = has_pc ? add(cur, sub(cur, pc)) : cur;
                const vec2 c2 = P(g0, g1), to = P(g2, g3);
                cur = to;
                pc = c2;
                has_pc = true;
                O.cubic_to(c1, c2, to);
                break;
            }
            case 'a': {
                const vec2 from = cur;
                const vec2 to = P(g5, g6);
                cur = to;
                // Angle::degrees: f32::to_radians, x * (PI / 180)
                const float xrot = g2 * (3.14159265358979323846f / 180.0f);
                // SvgArc::is_straight_line (arc.rs:534-538)
                const float EPS = 1.1920929e-07f;
                if (ffabs(g0) <= EPS || ffabs(g1) <= EPS || veq(from, to)) {
                    O.line_to(to);
                } else {
                    const ArcD arc = svg_arc_to_arc(from, to, mk(g0, g1), xrot, g3 != 0.0f,
                                                    g4 != 0.0f);
                    arc_to_quads_each(arc, [&](int, vec2, vec2 ctrl, vec2 qto) {
                        O.quad_to(ctrl, qto);
                    });
                }
                break;
            }
            default: {  // m
                const vec2 to = P(g0, g1);
                cur = to;
                first_pos = to;
                O.begin(to);
                need_end = true;
                need_start = false;
                break;
            }
        }
        const uint32_t lc = cmd | 0x20u;
        if (lc == 'c' || lc == 's') {
            has_pq = false;
        } else if (lc == 'q' || lc == 't') {
            has_pc = false;
        } else {
            has_pc = has_pq = false;
        }
        implicit = cmd == 'm' ? 'l' : cmd == 'M' ? 'L' : cmd;
        skip_ws(S);
    }
    if (need_end) O.end(false);
    if (S.non_ascii) E = PErr{LYG_SVG_ERR_NON_ASCII, S.na_line, S.na_col, 0u};
}

}  // namespace

// Parses string i into its staging room: verbs at text_off[i] + 2 i (room: its bytes + 2),
// points at 2 text_off[i] + 2 i (room: twice its bytes + 2).  Every command consumes at
// least one character and makes at most one verb and two points per character (an arc's
// eight quadratics need eleven), plus the End of the last sub-path; a string whose output
// would not fit anyway raises the overflow counter instead of writing past its room.
__global__ void __launch_bounds__(256) svg_parse_kernel(SvgLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const uint32_t a = L.text_off[i], b = L.text_off[i + 1];
    PErr E;
    Out<true> O{L.verbs + a + 2ull * i, reinterpret_cast<float2*>(L.points) + 2ull * a + 2ull * i,
                0, 0, mk(0.0f, 0.0f), b - a + 2, 2 * (b - a) + 2};
    parse_svg<true>(L.text, a, b, L.num_attributes, L.stop_at, O, E);
    L.counts_v[i] = O.nv;
    L.counts_p[i] = O.np;
    if (O.nv > O.cap_v || O.np > O.cap_p) atomicAdd(L.n_errors + 1, 1u);
    if (L.errors) {
        int32_t* e = L.errors + 4ull * i;
        e[0] = E.kind;
        e[1] = E.line;
        e[2] = E.col;
        e[3] = (int32_t)E.ch;
    }
    if (E.kind != LYG_SVG_OK) atomicAdd(L.n_errors, 1u);
}

// Copies every string's verbs and points from its staging room to its place in the batch:
// one block per string.
__global__ void __launch_bounds__(256) svg_compact_kernel(SvgLaunch L, uint8_t* verbs,
                                                          float* points) {
    for (uint32_t i = blockIdx.x; i < L.n; i += gridDim.x) {
        const uint32_t a = L.text_off[i];
        const uint8_t* sv = L.verbs + a + 2ull * i;
        const float2* sp = reinterpret_cast<const float2*>(L.points) + 2ull * a + 2ull * i;
        const uint32_t ov = L.off_v[i], nv = L.off_v[i + 1] - ov;
        const uint32_t op = L.off_p[i], np = L.off_p[i + 1] - op;
        for (uint32_t k = threadIdx.x; k < nv; k += blockDim.x) verbs[ov + k] = sv[k];
        float2* dp = reinterpret_cast<float2*>(points) + op;
        for (uint32_t k = threadIdx.x; k < np; k += blockDim.x) dp[k] = sp[k];
    }
}

}  // namespace lyg

extern "C" {

hipError_t lyg_launch_svg_parse(const lyg::SvgLaunch* L, hipStream_t s) {
    if (L->n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::svg_parse_kernel, dim3((L->n + 255) / 256), dim3(256), 0, s, *L);
    return hipGetLastError();
}

hipError_t lyg_launch_svg_compact(const lyg::SvgLaunch* L, uint8_t* verbs, float* points,
                                  hipStream_t s) {
    if (L->n == 0) return hipSuccess;
    const uint32_t blocks = L->n < 65536u ? L->n : 65536u;
    hipLaunchKernelGGL(lyg::svg_compact_kernel, dim3(blocks), dim3(256), 0, s, *L, verbs, points);
    return hipGetLastError();
}

}  // extern "C"
