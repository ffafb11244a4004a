// Seeded synthetic path generators for the benchmark workloads.
//
// SplitMix64 seeds one PCG32 (XSH-RR) stream per path: path i of a batch with seed s
// uses pcg(splitmix64(s ^ splitmix64(i + 1))), so paths can be generated in any order
// and every batch is byte-stable across runs and machines.  Uniform floats are
// (u32 >> 8) * 2^-24 in [0, 1).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/lyg.h"

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
    const double lmin = std::log((double)min_segs), lmax = std::log((double)max_segs);
    for (uint32_t i = 0; i < n_paths; ++i) {
        Pcg32 rng(path_seed(seed, i));
        double u = rng.uniform();
        double segs_d = std::round(std::exp(lmin + u * (lmax - lmin)));
        uint32_t segs = (uint32_t)segs_d;
        if (segs < min_segs) segs = min_segs;
        if (segs > max_segs) segs = max_segs;
        if (verb_offsets) verb_offsets[i] = (uint32_t)nv;
        if (point_offsets) point_offsets[i] = (uint32_t)np;
        auto pt = [&]() {
            float x = rng.uniform() * coord_max;
            float y = rng.uniform() * coord_max;
            if (points) {
                points[2 * np] = x;
                points[2 * np + 1] = y;
            }
            np++;
        };
        auto vb = [&](uint8_t v) {
            if (verbs) verbs[nv] = v;
            nv++;
        };
        vb(LYG_VERB_BEGIN);
        pt();
        for (uint32_t s = 0; s < segs; ++s) {
            float k = rng.uniform();
            if (k < 0.25f) {
                vb(LYG_VERB_LINE);
                pt();
            } else if (k < 0.5f) {
                vb(LYG_VERB_QUADRATIC);
                pt();
                pt();
            } else {
                vb(LYG_VERB_CUBIC);
                pt();
                pt();
                pt();
            }
        }
        if (closed) {
            // BuilderImpl::end(true) stores a copy of the first point (path.rs:537-542).
            vb(LYG_VERB_CLOSE);
            if (points) {
                uint32_t first = point_offsets ? point_offsets[i] : 0;
                points[2 * np] = points[2 * first];
                points[2 * np + 1] = points[2 * first + 1];
            }
            np++;
        } else {
            vb(LYG_VERB_END);
        }
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

}  // extern "C"
