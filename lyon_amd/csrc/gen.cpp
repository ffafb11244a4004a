// Seeded synthetic path generators for the benchmark workloads.
//
// SplitMix64 seeds one PCG32 (XSH-RR) stream per path: path i of a batch with seed s
// uses pcg(splitmix64(s ^ splitmix64(i + 1))), so paths can be generated in any order
// and every batch is byte-stable across runs and machines.  Uniform floats are
// (u32 >> 8) * 2^-24 in [0, 1).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/lyg.h"
#include "lyg_math.h"

namespace {

uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Pcg32 {
    uint64_t state, inc;
    explicit Pcg32(uint64_t seed, uint64_t stream = 0xda3e39cb94b95bdbull) {
        state = 0;
        inc = (stream << 1) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
};

uint64_t path_seed(uint64_t seed, uint64_t i) { return splitmix64(seed ^ splitmix64(i + 1)); }

}  // namespace

// One path of each workload, appended to `verbs` / `points` (x, y pairs).
namespace lygen {

void stroke_path(uint64_t seed, uint64_t i, uint32_t min_segs, uint32_t max_segs,
                 float coord_max, bool closed, std::vector<uint8_t>& verbs,
                 std::vector<float>& points) {
    Pcg32 rng(path_seed(seed, i));
    const double lmin = std::log((double)min_segs), lmax = std::log((double)max_segs);
    double u = rng.uniform();
    uint32_t segs = (uint32_t)std::round(std::exp(lmin + u * (lmax - lmin)));
    segs = segs < min_segs ? min_segs : (segs > max_segs ? max_segs : segs);
    const size_t first = points.size();
    auto pt = [&]() {
        float x = rng.uniform() * coord_max;
        float y = rng.uniform() * coord_max;
        points.push_back(x);
        points.push_back(y);
    };
    verbs.push_back(LYG_VERB_BEGIN);
    pt();
    for (uint32_t s = 0; s < segs; ++s) {
        float k = rng.uniform();
        if (k < 0.25f) {
            verbs.push_back(LYG_VERB_LINE);
            pt();
        } else if (k < 0.5f) {
            verbs.push_back(LYG_VERB_QUADRATIC);
            pt();
            pt();
        } else {
            verbs.push_back(LYG_VERB_CUBIC);
            pt();
            pt();
            pt();
        }
    }
    if (closed) {
        // BuilderImpl::end(true) stores a copy of the first point (path.rs:537-542).
        verbs.push_back(LYG_VERB_CLOSE);
        float fx = points[first], fy = points[first + 1];
        points.push_back(fx);
        points.push_back(fy);
    } else {
        verbs.push_back(LYG_VERB_END);
    }
}

void fill_path(uint64_t seed, uint64_t i, std::vector<uint8_t>& verbs,
               std::vector<float>& pts) {
    const double two_pi = 6.283185307179586;
    Pcg32 rng(path_seed(seed ^ 0x5eedf111ull, i));
    auto vb = [&](uint8_t v) { verbs.push_back(v); };
    const float R = 16.0f + rng.uniform() * (512.0f - 16.0f);
    const float cx = R + rng.uniform() * (4096.0f - 2.0f * R);
    const float cy = R + rng.uniform() * (4096.0f - 2.0f * R);
    const size_t first = pts.size();
    if ((i & 1u) == 0) {
        uint32_t n = 5u + rng.next() % 508u;
        vb(LYG_VERB_BEGIN);
        for (uint32_t j = 0; j < n; ++j) {
            double a = two_pi * rng.uniform();
            float r = R * (0.2f + 0.8f * rng.uniform());
            pts.push_back(cx + r * (float)std::cos(a));
            pts.push_back(cy + r * (float)std::sin(a));
            if (j > 0) vb(LYG_VERB_LINE);
        }
    } else {
        uint32_t m = 4u + rng.next() % 253u;
        vb(LYG_VERB_BEGIN);
        std::vector<float> ax(m), ay(m);
        for (uint32_t j = 0; j < m; ++j) {
            double a = two_pi * (double)j / (double)m;
            float r = R * (0.5f + 0.5f * rng.uniform());
            ax[j] = cx + r * (float)std::cos(a);
            ay[j] = cy + r * (float)std::sin(a);
        }
        pts.push_back(ax[0]);
        pts.push_back(ay[0]);
        for (uint32_t j = 0; j < m; ++j) {
            uint32_t q = (j + 1) % m;
            for (int c = 0; c < 2; ++c) {
                float bx = c == 0 ? ax[j] : ax[q], by = c == 0 ? ay[j] : ay[q];
                double d = two_pi * rng.uniform();
                float len = R * 0.6f * rng.uniform();
                pts.push_back(bx + len * (float)std::cos(d));
                pts.push_back(by + len * (float)std::sin(d));
            }
            pts.push_back(ax[q]);
            pts.push_back(ay[q]);
            vb(LYG_VERB_CUBIC);
        }
        // scale about the centre until 16 <= flattened edges <= 2048
        for (int it = 0; it < 32; ++it) {
            uint64_t f = 0;
            for (uint32_t j = 0; j < m; ++j) {
                const float* p = &pts[first + 6 * j];
                lyg::vec2 a = lyg::mk(p[0], p[1]), b = lyg::mk(p[2], p[3]);
                lyg::vec2 c = lyg::mk(p[4], p[5]), d = lyg::mk(p[6], p[7]);
                f += lyg::seg_count(lyg::cubic_wang(a, b, c, d, 0.05f));
            }
            float k;
            if (f > 2048) k = 0.8f * (float)(2048.0 / (double)f) * (float)(2048.0 / (double)f);
            else if (f < 16) k = 1.25f * (float)(16.0 / (double)f) * (float)(16.0 / (double)f);
            else break;
            for (size_t e = first; e < pts.size(); e += 2) {
                pts[e] = cx + (pts[e] - cx) * k;
                pts[e + 1] = cy + (pts[e + 1] - cy) * k;
            }
        }
        for (size_t e = first; e < pts.size(); ++e)
            pts[e] = std::fmin(std::fmax(pts[e], 0.0f), 4096.0f);
    }
    // BuilderImpl::end(true) stores a copy of the first point (path.rs:537-542).
    vb(LYG_VERB_CLOSE);
    float fx = pts[first], fy = pts[first + 1];
    pts.push_back(fx);
    pts.push_back(fy);
}

}  // namespace lygen

namespace lygen {
// Random walk of n vertices, stream `stream` of `seed` (see lyg_gen_random_walk).
void random_walk(uint64_t seed, uint64_t stream, uint32_t n, float extent, uint8_t* verbs,
                 float* points) {
    const double two_pi = 6.283185307179586;
    Pcg32 rng(path_seed(seed ^ 0x7a1c3a1cull, stream));
    float x = 0.0f, y = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        points[2 * i] = x;
        points[2 * i + 1] = y;
        const double u1 = ((double)(rng.next() >> 8) + 1.0) * (1.0 / 16777216.0);  // (0, 1]
        const double u2 = (double)(rng.next() >> 8) * (1.0 / 16777216.0);
        const double r = std::sqrt(-2.0 * std::log(u1));
        x += (float)(r * std::cos(two_pi * u2));
        y += (float)(r * std::sin(two_pi * u2));
    }
    float mn[2] = {points[0], points[1]}, mx[2] = {points[0], points[1]};
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 2; ++a) {
            mn[a] = std::fmin(mn[a], points[2 * i + a]);
            mx[a] = std::fmax(mx[a], points[2 * i + a]);
        }
    for (uint32_t i = 0; i < n; ++i)
        for (int a = 0; a < 2; ++a) {
            const double span = (double)mx[a] - (double)mn[a];
            const double t = span > 0 ? ((double)points[2 * i + a] - (double)mn[a]) / span : 0.0;
            points[2 * i + a] = (float)(t * (double)extent);
        }
    verbs[0] = LYG_VERB_BEGIN;
    for (uint32_t i = 1; i < n; ++i) verbs[i] = LYG_VERB_LINE;
    verbs[n] = LYG_VERB_CLOSE;
    points[2 * n] = points[0];
    points[2 * n + 1] = points[1];
}

// Fill request i of the serving stream: a closed random walk whose edge count is drawn
// log-uniformly from [min_segs, max_segs] (as stroke_path draws its segment count),
// appended to `verbs` / `points`.
void walk_path(uint64_t seed, uint64_t i, uint32_t min_segs, uint32_t max_segs, float extent,
               std::vector<uint8_t>& verbs, std::vector<float>& points) {
    Pcg32 rng(path_seed(seed ^ 0x3a1c3a1cull, i));
    const double lmin = std::log((double)min_segs), lmax = std::log((double)max_segs);
    uint32_t segs = (uint32_t)std::round(std::exp(lmin + rng.uniform() * (lmax - lmin)));
    segs = segs < min_segs ? min_segs : (segs > max_segs ? max_segs : segs);
    if (segs < 3) segs = 3;
    const size_t v0 = verbs.size(), p0 = points.size();
    verbs.resize(v0 + segs + 1);
    points.resize(p0 + 2 * (segs + 1));
    random_walk(seed, i + 1, segs, extent, verbs.data() + v0, points.data() + p0);
}
}  // namespace lygen

extern "C" {

// Open (or closed) paths with a log-uniform segment count in [min_segs, max_segs],
// segment kinds 25% line / 25% quadratic / 50% cubic, every point uniform in
// [0, coord_max]^2.  Two-call protocol: sizes first (null arrays), then the data.
int32_t lyg_gen_stroke_batch(uint64_t seed, uint32_t n_paths, uint32_t min_segs,
                             uint32_t max_segs, float coord_max, int closed, uint8_t* verbs,
                             float* points, uint32_t* verb_offsets, uint32_t* point_offsets,
                             uint64_t* n_verbs, uint64_t* n_points) {
    if (min_segs < 1 || max_segs < min_segs) return LYG_ERR_INVALID_ARGUMENT;
    uint64_t nv = 0, np = 0;
    std::vector<uint8_t> v;
    std::vector<float> p;
    for (uint32_t i = 0; i < n_paths; ++i) {
        v.clear();
        p.clear();
        lygen::stroke_path(seed, i, min_segs, max_segs, coord_max, closed != 0, v, p);
        if (verb_offsets) verb_offsets[i] = (uint32_t)nv;
        if (point_offsets) point_offsets[i] = (uint32_t)np;
        if (verbs) std::memcpy(verbs + nv, v.data(), v.size());
        if (points) std::memcpy(points + 2 * np, p.data(), 4 * p.size());
        nv += v.size();
        np += p.size() / 2;
    }
    if (verb_offsets) verb_offsets[n_paths] = (uint32_t)nv;
    if (point_offsets) point_offsets[n_paths] = (uint32_t)np;
    if (n_verbs) *n_verbs = nv;
    if (n_points) *n_points = np;
    return LYG_OK;
}

// Per-path stroke options of the batched stroke workload: width U[w_min, w_max],
// join uniform over {Miter, Round, Bevel}, one cap uniform over {Butt, Square, Round}
// for both ends, miter limit 4.
int32_t lyg_gen_stroke_options(uint64_t seed, uint32_t n_paths, float tolerance, float w_min,
                               float w_max, lyg_stroke_options* out) {
    static const uint8_t joins[3] = {LYG_JOIN_MITER, LYG_JOIN_ROUND, LYG_JOIN_BEVEL};
    for (uint32_t i = 0; i < n_paths; ++i) {
        Pcg32 rng(path_seed(seed ^ 0x0123456789abcdefull, i));
        lyg_stroke_options o;
        o.line_width = w_min + rng.uniform() * (w_max - w_min);
        o.miter_limit = 4.0f;
        o.tolerance = tolerance;
        uint32_t j = rng.next() % 3u;
        uint32_t c = rng.next() % 3u;
        o.line_join = joins[j];
        o.start_cap = (uint8_t)c;
        o.end_cap = (uint8_t)c;
        o._pad = 0;
        out[i] = o;
    }
    return LYG_OK;
}

// The batched fill workload: closed paths in [0, 4096]^2, even paths are stars, odd
// paths cubic loops.
//  * star: n points, n uniform in [5, 512], at angles uniform in [0, 2*pi) in draw order
//    (unsorted, so the polygon crosses itself), radius R * U[0.2, 1] with R uniform in
//    [16, 512], around a centre in [R, 4096 - R]^2;
//  * cubic loop: m cubics, m uniform in [4, 256], through anchors at angles 2*pi*j/m and
//    radius R * U[0.5, 1], control points at anchor + R * U[0, 0.6] in a random direction
//    (loops and cusps appear); the loop is scaled about its centre until it flattens to
//    16-2048 edges at tolerance 0.05 (Wang's counts, as the fill's event builder).
int32_t lyg_gen_fill_batch(uint64_t seed, uint32_t n_paths, uint8_t* verbs, float* points,
                           uint32_t* verb_offsets, uint32_t* point_offsets, uint64_t* n_verbs,
                           uint64_t* n_points) {
    uint64_t nv = 0, np = 0;
    std::vector<uint8_t> v;
    std::vector<float> p;
    for (uint32_t i = 0; i < n_paths; ++i) {
        v.clear();
        p.clear();
        lygen::fill_path(seed, i, v, p);
        if (verb_offsets) verb_offsets[i] = (uint32_t)nv;
        if (point_offsets) point_offsets[i] = (uint32_t)np;
        if (verbs) std::memcpy(verbs + nv, v.data(), v.size());
        if (points) std::memcpy(points + 2 * np, p.data(), 4 * p.size());
        nv += v.size();
        np += p.size() / 2;
    }
    if (verb_offsets) verb_offsets[n_paths] = (uint32_t)nv;
    if (point_offsets) point_offsets[n_paths] = (uint32_t)np;
    if (n_verbs) *n_verbs = nv;
    if (n_points) *n_points = np;
    return LYG_OK;
}

// The single large fill workload: one closed polyline of n vertices from a 2D random walk
// (each step N(0, 1) per axis by Box-Muller on PCG32, accumulated in f32), then rescaled
// per axis (f64, rounded to f32) onto [0, extent]^2.  Begin, n - 1 Lines and Close (with
// the stored copy of the first point): n + 1 verbs, n + 1 points.

int32_t lyg_gen_random_walk(uint64_t seed, uint32_t n, float extent, uint8_t* verbs,
                            float* points) {
    if (n < 2 || !verbs || !points) return LYG_ERR_INVALID_ARGUMENT;
    lygen::random_walk(seed, 0, n, extent, verbs, points);
    return LYG_OK;
}

// Per-path FillOptions of the batched fill workload: EvenOdd / NonZero with equal
// probability, vertical sweep, intersections handled.
int32_t lyg_gen_fill_options(uint64_t seed, uint32_t n_paths, float tolerance,
                             lyg_fill_options* out) {
    for (uint32_t i = 0; i < n_paths; ++i) {
        Pcg32 rng(path_seed(seed ^ 0x0f111f111ull, i));
        lyg_fill_options o;
        o.tolerance = tolerance;
        o.fill_rule = (uint8_t)(rng.next() & 1u);
        o.sweep_orientation = LYG_ORIENT_VERTICAL;
        o.handle_intersections = 1;
        o._pad = 0;
        out[i] = o;
    }
    return LYG_OK;
}

}  // extern "C"
