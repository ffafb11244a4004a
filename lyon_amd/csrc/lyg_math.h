// Shared scalar/vector arithmetic for the HIP kernels (also compiles as plain C++ so the
// host-side unit tests can exercise it).
//
// Everything here keeps lyon's f32 operation order so that results are bit-identical to
// the reference's Rust code under IEEE-754 (the kernels are built with
// -ffp-contract=off and correctly rounded f32 divide/sqrt):
//   * vector helpers follow euclid 0.22.9's Vector2D/Point2D arithmetic,
//   * wang_count follows quadratic_bezier.rs:943-974 / cubic_bezier.rs:1369-1403,
//   * tchain reproduces the serial `t0 + step` accumulation of
//     for_each_flattened_with_t (quadratic_bezier.rs:377-386, cubic_bezier.rs:599-611)
//     in O(log n) integer steps so that any lane can start at any flattened point.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LYG_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#include <cstring>
#define LYG_HD inline
#endif

namespace lyg {

struct vec2 {
    float x, y;
};
LYG_HD vec2 mk(float x, float y) {
    vec2 r;
    r.x = x;
    r.y = y;
    return r;
}
LYG_HD vec2 add(vec2 a, vec2 b) { return mk(a.x + b.x, a.y + b.y); }
LYG_HD vec2 sub(vec2 a, vec2 b) { return mk(a.x - b.x, a.y - b.y); }
LYG_HD vec2 neg(vec2 a) { return mk(-a.x, -a.y); }
LYG_HD vec2 mul(vec2 a, float s) { return mk(a.x * s, a.y * s); }
LYG_HD vec2 divs(vec2 a, float s) { return mk(a.x / s, a.y / s); }
LYG_HD bool veq(vec2 a, vec2 b) { return a.x == b.x && a.y == b.y; }
LYG_HD float dot(vec2 a, vec2 b) { return a.x * b.x + a.y * b.y; }
LYG_HD float cross(vec2 a, vec2 b) { return a.x * b.y - a.y * b.x; }
LYG_HD float sqlen(vec2 a) { return a.x * a.x + a.y * a.y; }

#if defined(__HIPCC__)
LYG_HD float fsqrt(float v) { return __builtin_sqrtf(v); }
LYG_HD float ffabs(float v) { return __builtin_fabsf(v); }
LYG_HD float ffmax(float a, float b) { return __builtin_fmaxf(a, b); }
LYG_HD float ffmin(float a, float b) { return __builtin_fminf(a, b); }
LYG_HD uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
LYG_HD float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
#else
LYG_HD float fsqrt(float v) { return std::sqrt(v); }
LYG_HD float ffabs(float v) { return std::fabs(v); }
LYG_HD float ffmax(float a, float b) { return std::fmax(a, b); }
LYG_HD float ffmin(float a, float b) { return std::fmin(a, b); }
LYG_HD uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
LYG_HD float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
#endif

LYG_HD float vlen(vec2 a) { return fsqrt(sqlen(a)); }
LYG_HD vec2 normalize(vec2 a) { return divs(a, vlen(a)); }
LYG_HD vec2 perp(vec2 v) { return mk(-v.y, v.x); }  // lyon_geom utils::tangent
LYG_HD vec2 lerp(vec2 a, vec2 b, float t) {
    float one_t = 1.0f - t;
    return mk(one_t * a.x + t * b.x, one_t * a.y + t * b.y);
}

// --------------------------------------------------------------------------------------
// Flattening segment counts (Wang's formula with lyon's i^4 lookup table).
// --------------------------------------------------------------------------------------
LYG_HD float wang_tail(float err4) {
    if (err4 <= 331776.0f) {
        // Smallest i in 1..24 with err4 <= i^4; i^4 is exact in f32 for these i.
        for (int i = 1; i <= 24; ++i) {
            float i2 = (float)(i * i);
            if (err4 <= i2 * i2) return (float)i;
        }
    }
    return ffmax(fsqrt(fsqrt(err4)), 1.0f);
}
LYG_HD float quad_wang(vec2 from, vec2 ctrl, vec2 to, float tol) {
    vec2 l = mul(add(sub(from, mul(ctrl, 2.0f)), to), 2.0f);
    float d = 1.0f / (8.0f * tol);
    float err4 = dot(l, l) * d * d;
    return wang_tail(err4);
}
LYG_HD float cubic_wang(vec2 from, vec2 c1, vec2 c2, vec2 to, float tol) {
    vec2 v1 = mul(add(sub(from, mul(c1, 2.0f)), c2), 6.0f);
    vec2 v2 = mul(add(sub(c1, mul(c2, 2.0f)), to), 6.0f);
    float l = ffmax(dot(v1, v1), dot(v2, v2));
    float d = 1.0f / (8.0f * tol);
    float err4 = l * d * d;
    return wang_tail(err4);
}
// ToPrimitive::to_u32(n).unwrap_or(1) for the f32 segment count.
LYG_HD uint32_t seg_count(float n) {
    if (!(n > -1.0f) || !(n < 4294967296.0f)) return 1u;
    return (uint32_t)n;
}

// --------------------------------------------------------------------------------------
// tchain(s, j): the f32 value of t after j serial additions `t = t + s` from t = 0.
//
// With s = m * q (m the 24-bit significand, q = ulp(s)) every partial sum is an integer
// multiple T of q.  A step rounds S = T + m to the f32 grid of S's binade, whose spacing
// in units of q is g = 2^b with b = bitlen(S) - 24.  Inside one binade, once T is a
// multiple of g, every step adds the same amount: M*g or (M+1)*g (M = m >> b) depending
// on how the remainder r = m mod g compares with g/2; on an exact tie the increment is
// the even one of {M, M+1} as soon as T/g is even (round-half-to-even), which one step
// guarantees.  So the chain is a short list of linear runs, one or two per binade, and
// j steps cost O(log j) integer operations.  s must be a positive normal float.
// --------------------------------------------------------------------------------------
LYG_HD int bitlen64(uint64_t v) {
#if defined(__HIPCC__)
    return v ? 64 - __builtin_clzll(v) : 0;
#else
    return v ? 64 - __builtin_clzll(v) : 0;
#endif
}
// Round S (units of q) to the f32 grid of its binade, ties to even.
LYG_HD uint64_t tchain_round(uint64_t S) {
    int b = bitlen64(S) - 24;
    if (b <= 0) return S;
    uint64_t g = 1ull << b;
    uint64_t A = S >> b;
    uint64_t r = S & (g - 1);
    uint64_t half = g >> 1;
    if (r > half || (r == half && (A & 1ull))) A += 1;
    return A << b;
}
// T after j steps, continuing from a known state: T0 = T after k0 steps (0, 0 from the
// start).  The loop only reads (T, k), so any earlier state of the same chain is a valid
// start.
LYG_HD uint64_t tchain_units_from(uint32_t m, uint32_t j, uint32_t k0, uint64_t T0) {
    uint64_t T = T0;
    uint32_t k = k0;
    while (k < j) {
        uint64_t S = T + m;
        int b = bitlen64(S) - 24;
        if (b <= 0) {
            // Exact region: all sums below 2^24 units are representable.  Here S < 2^24,
            // so the step count is a 32-bit division (64-bit division is a long
            // software routine on the GPU).
            const uint32_t lim = (1u << 24) - 1;  // largest exact sum
            // number of further exact steps: T + c*m <= lim
            uint32_t c = (lim - (uint32_t)T) / m;
            if (c == 0) c = 1;
            const uint32_t left = j - k;
            if (c > left) c = left;
            T += (uint64_t)c * m;
            k += c;
            continue;
        }
        uint64_t g = 1ull << b;
        uint64_t A = T >> b;
        bool aligned = (T & (g - 1)) == 0;
        uint64_t M = (uint64_t)m >> b;
        uint64_t r = (uint64_t)m & (g - 1);
        uint64_t half = g >> 1;
        if (!aligned || (r == half && (A & 1ull))) {
            T = tchain_round(S);
            k += 1;
            continue;
        }
        uint64_t inc_units;
        if (r < half) inc_units = M;
        else if (r > half) inc_units = M + 1;
        else inc_units = (M & 1ull) ? M + 1 : M;
        if (inc_units == 0) break;  // step below half an ulp: t no longer moves
        uint64_t inc = inc_units << b;
        // Steps whose sum stays below the binade end 2^(24+b): T + i*inc + m < lim_end.
        uint64_t lim_end = 1ull << (24 + b);
        uint64_t room = lim_end - 1 - S;  // S = T + m < lim_end
        // room / inc = (room >> b) / inc_units (floor of floors), with room < 2^(23+b)
        // (S >= 2^(23+b)) and inc_units < 2^24: a 32-bit division.
        uint32_t c = (uint32_t)(room >> b) / (uint32_t)inc_units + 1;
        const uint32_t left = j - k;
        if (c > left) c = left;
        T += (uint64_t)c * inc;
        k += c;
    }
    return T;
}
LYG_HD uint64_t tchain_units(uint32_t m, uint32_t j) { return tchain_units_from(m, j, 0, 0); }
// The 24-bit significand m of s (s = m * ulp(s)).
LYG_HD uint32_t tchain_mant(float s) { return (f2u(s) & 0x7fffffu) | 0x800000u; }
// The f32 value T * ulp(s).
LYG_HD float tchain_value(float s, uint64_t T) {
    uint32_t e = (f2u(s) >> 23) & 0xffu;
    // value = T * 2^(e - 150); T < 2^(24 + b) is a multiple of 2^b, so T >> b fits in 24
    // bits and the conversion below is exact.
    int b = bitlen64(T) - 24;
    if (b < 0) b = 0;
    uint32_t A = (uint32_t)(T >> b);
    int ex = (int)e - 150 + b;  // value = A * 2^ex with A < 2^24
    // Build A * 2^ex exactly: A as float (exact), then scale by a power of two.
    float fa = (float)A;
    // ex is in [-150, 0] for t <= ~1; scale in two steps to stay in the normal range.
    int e1 = ex / 2, e2 = ex - e1;
    float s1 = u2f((uint32_t)(e1 + 127) << 23);
    float s2 = u2f((uint32_t)(e2 + 127) << 23);
    return (fa * s1) * s2;
}
// The units T of a value t of the chain (t = T * ulp(s), t >= s > 0 normal): the exact
// inverse of tchain_value.
LYG_HD uint64_t tchain_units_of(float s, float t) {
    const uint32_t es = (f2u(s) >> 23) & 0xffu, bt = f2u(t);
    const uint32_t et = (bt >> 23) & 0xffu;
    return (uint64_t)((bt & 0x7fffffu) | 0x800000u) << (et - es);
}
LYG_HD float tchain(float s, uint32_t j) {
    if (j == 0) return 0.0f;
    return tchain_value(s, tchain_units(tchain_mant(s), j));
}

}  // namespace lyg
