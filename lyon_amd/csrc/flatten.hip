// lyon_geom flattening as array kernels for gfx950.
//
// Segments (LineSegment / QuadraticBezierSegment / CubicBezierSegment): a count kernel
// evaluates Wang's formula per segment (one lane each), the host-side runtime scans the
// counts, and an emit kernel gives every output point its own lane: the lane finds its
// segment by binary search over the offsets, recovers t with the closed-form serial
// chain (lyg_math.h: tchain) and evaluates the polynomial form with Horner's rule in
// lyon's operation order (quadratic_bezier.rs:891-938, cubic_bezier.rs:1318-1364).
// Output is fully coalesced: consecutive lanes write consecutive points.
//
// Arcs (Arc::for_each_flattened_with_t, arc.rs:339-368) re-split the remaining arc at
// every step, a serial recurrence, so one lane walks one arc.
#include <hip/hip_runtime.h>

#include "lyg_math.h"
#include "lyg_types.h"
#include "stroke_core.h"

namespace lyg {

struct SegPoly {
    vec2 a0, a1, a2, a3;
    vec2 to;
    float step;
    uint32_t n;
    bool cubic;
};

__device__ __forceinline__ SegPoly seg_poly(int kind, const float* c, float tol) {
    SegPoly s;
    vec2 p0 = mk(c[0], c[1]), p1 = mk(c[2], c[3]), p2 = mk(c[4], c[5]), p3 = mk(c[6], c[7]);
    s.a0 = s.a1 = s.a2 = s.a3 = mk(0.0f, 0.0f);
    s.cubic = false;
    if (kind == 0) {
        s.n = 1;
        s.to = p1;
        s.step = 1.0f;
    } else if (kind == 1) {
        float n = quad_wang(p0, p1, p2, tol);
        s.step = 1.0f / n;
        s.n = seg_count(n);
        s.a0 = p0;
        s.a1 = mul(sub(p1, p0), 2.0f);
        s.a2 = sub(add(p0, p2), mul(p1, 2.0f));
        s.to = p2;
    } else {
        float n = cubic_wang(p0, p1, p2, p3, tol);
        s.step = 1.0f / n;
        s.n = seg_count(n);
        s.a0 = p0;
        s.a1 = mul(sub(p1, p0), 3.0f);
        s.a2 = add(sub(mul(p0, 3.0f), mul(p1, 6.0f)), mul(p2, 3.0f));
        s.a3 = add(sub(p3, p0), mul(sub(p1, p2), 3.0f));
        s.to = p3;
        s.cubic = true;
    }
    return s;
}

__global__ void flatten_count_kernel(FlattenLaunch L) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    SegPoly s = seg_poly(L.kinds[i], L.ctrl + 8ull * i, L.tolerance);
    L.counts[i] = s.n;
}

__global__ void flatten_emit_kernel(FlattenLaunch L, uint32_t total) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    // last segment whose offset <= g
    uint32_t lo = 0, hi = L.n - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (L.offsets[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    SegPoly s = seg_poly(L.kinds[lo], L.ctrl + 8ull * lo, L.tolerance);
    uint32_t k = g - L.offsets[lo];
    vec2 p;
    if (k + 1 >= s.n) {
        p = s.to;
    } else {
        float t = tchain(s.step, k + 1);
        vec2 v = s.a0;
        float tt = t;
        v = add(v, mul(s.a1, tt));
        tt = tt * t;
        v = add(v, mul(s.a2, tt));
        if (s.cubic) {
            tt = tt * t;
            v = add(v, mul(s.a3, tt));
        }
        p = v;
    }
    reinterpret_cast<float2*>(L.out_points)[g] = make_float2(p.x, p.y);
}

// ---- arcs -------------------------------------------------------------------------------
struct ArcD {
    vec2 center, radii;
    float start, sweep, xrot;
};
__device__ __forceinline__ vec2 arc_sample(const ArcD& a, float t) {
    float angle = a.start + a.sweep * t;
    vec2 e = mk(a.radii.x * t_cos(angle), a.radii.y * t_sin(angle));
    return add(a.center, rotate_v(a.xrot, e));
}
__device__ __forceinline__ float arc_step(const ArcD& a, float tol) {
    float r = vlen(sub(arc_sample(a, 0.0f), a.center));
    float ang = 2.0f * t_acos((r - tol) / r);
    float res = ffmin(ang / ffabs(a.sweep), 1.0f);
    if (res < 1e-4f) return 1.0f;
    return res;
}
// Walks one arc; writes points when `out` is non-null, returns the count.
__device__ uint32_t arc_walk(ArcD arc, float tol, float2* out) {
    ArcD it = arc;
    uint32_t n = 0;
    for (;;) {
        float step = arc_step(it, tol);
        if (step >= 1.0f) break;
        float split = it.sweep * step;
        it.start = it.start + split;
        it.sweep = it.sweep - split;
        vec2 to = arc_sample(it, 0.0f);
        if (out) out[n] = make_float2(to.x, to.y);
        ++n;
        if (n > (1u << 24)) break;  // runaway guard (tolerance far below the radius' ulp)
    }
    vec2 last = arc_sample(arc, 1.0f);
    if (out) out[n] = make_float2(last.x, last.y);
    return n + 1;
}
__global__ void arc_flatten_kernel(const float* arcs, uint32_t n, float tol, uint32_t* counts,
                                   const uint32_t* offsets, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = arcs + 7ull * i;
    ArcD arc{mk(a[0], a[1]), mk(a[2], a[3]), a[4], a[5], a[6]};
    if (counts) counts[i] = arc_walk(arc, tol, nullptr);
    else arc_walk(arc, tol, reinterpret_cast<float2*>(out) + offsets[i]);
}

// arc_to_quadratic_beziers_with_t (arc.rs:668-701)
__global__ void arc_to_quads_kernel(const float* arcs, uint32_t n, uint32_t* counts,
                                    const uint32_t* offsets, float* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float PI = 3.14159265358979323846f;
    const float FRAC_PI_4 = 0.785398163397448309616f;
    const float* a = arcs + 7ull * i;
    ArcD arc{mk(a[0], a[1]), mk(a[2], a[3]), a[4], a[5], a[6]};
    float sign = __builtin_copysignf(1.0f, arc.sweep);
    if (__builtin_isnan(arc.sweep)) sign = arc.sweep;
    float sweep = ffmin(ffabs(arc.sweep), PI * 2.0f);
    float n_steps = __builtin_ceilf(sweep / FRAC_PI_4);
    int ns = (n_steps >= 0.0f && n_steps < 2147483647.0f) ? (int)n_steps : 0;
    if (counts) {
        counts[i] = (uint32_t)ns;
        return;
    }
    float step = sweep / n_steps * sign;
    float alpha = (float)tan((double)(step * 0.5f));
    float* o = out + 6ull * offsets[i];
    for (int k = 0; k < ns; ++k) {
        float a1 = arc.start + step * (float)k;
        float a2 = arc.start + step * (float)(k + 1);
        vec2 v1 = rotate_v(arc.xrot, mk(arc.radii.x * t_cos(a1), arc.radii.y * t_sin(a1)));
        vec2 v2 = rotate_v(arc.xrot, mk(arc.radii.x * t_cos(a2), arc.radii.y * t_sin(a2)));
        vec2 from = add(arc.center, v1);
        vec2 to = add(arc.center, v2);
        vec2 tan_ = rotate_v(arc.xrot, mk(-arc.radii.x * t_sin(a1), arc.radii.y * t_cos(a1)));
        vec2 ctrl = add(from, mul(tan_, alpha));
        o[6 * k + 0] = from.x;
        o[6 * k + 1] = from.y;
        o[6 * k + 2] = ctrl.x;
        o[6 * k + 3] = ctrl.y;
        o[6 * k + 4] = to.x;
        o[6 * k + 5] = to.y;
    }
}

// ---- exclusive scan of u32 counts into n+1 offsets (single workgroup) -----------------
// 1024 threads, each owning a contiguous slice; totals are carried in 64-bit and an
// overflow of the 32-bit offsets is reported through `overflow`.
__global__ void __launch_bounds__(1024) scan_offsets_kernel(const uint32_t* counts, uint32_t n,
                                                            uint32_t* offsets, uint64_t* total,
                                                            uint32_t* overflow) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t b = t * per, e = min(n, b + per);
    uint64_t s = 0;
    for (uint32_t i = b; i < e; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
        uint64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - s;
    bool ovf = false;
    for (uint32_t i = b; i < e; ++i) {
        if (run > 0xffffffffull) ovf = true;
        offsets[i] = (uint32_t)run;
        run += counts[i];
    }
    if (t == 1023) {
        if (part[1023] > 0xffffffffull) ovf = true;
        offsets[n] = (uint32_t)part[1023];
        *total = part[1023];
    }
    if (ovf) atomicOr(overflow, 1u);
}

__global__ void fill_u32_kernel(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Device-to-device copy, 16 bytes per thread per step (grid-stride); the tail is copied
// byte by byte by the first threads.
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16,
                            const uint8_t* __restrict__ src_b, uint8_t* __restrict__ dst_b,
                            uint64_t tail_from, uint64_t tail_n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < tail_n) dst_b[tail_from + t] = src_b[tail_from + t];
}

}  // namespace lyg

extern "C" {

hipError_t lyg_launch_copy(void* dst, const void* src, uint64_t bytes, int grid, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const uint64_t n16 = bytes / 16;
    hipLaunchKernelGGL(lyg::copy_kernel, dim3(grid), dim3(256), 0, s,
                       reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), n16,
                       reinterpret_cast<const uint8_t*>(src), reinterpret_cast<uint8_t*>(dst),
                       16 * n16, bytes - 16 * n16);
    return hipGetLastError();
}

hipError_t lyg_launch_flatten_count(const lyg::FlattenLaunch* L, hipStream_t s) {
    if (L->n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::flatten_count_kernel, dim3((L->n + 255) / 256), dim3(256), 0, s, *L);
    return hipGetLastError();
}
hipError_t lyg_launch_flatten_emit(const lyg::FlattenLaunch* L, uint32_t total, hipStream_t s) {
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::flatten_emit_kernel, dim3((total + 255) / 256), dim3(256), 0, s, *L,
                       total);
    return hipGetLastError();
}
hipError_t lyg_launch_arc_flatten(const float* arcs, uint32_t n, float tol, uint32_t* counts,
                                  const uint32_t* offsets, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::arc_flatten_kernel, dim3((n + 127) / 128), dim3(128), 0, s, arcs, n,
                       tol, counts, offsets, out);
    return hipGetLastError();
}
hipError_t lyg_launch_arc_to_quads(const float* arcs, uint32_t n, uint32_t* counts,
                                   const uint32_t* offsets, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::arc_to_quads_kernel, dim3((n + 127) / 128), dim3(128), 0, s, arcs, n,
                       counts, offsets, out);
    return hipGetLastError();
}
hipError_t lyg_launch_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total,
                           uint32_t* overflow, hipStream_t s) {
    hipLaunchKernelGGL(lyg::scan_offsets_kernel, dim3(1), dim3(1024), 0, s, counts, n, offsets,
                       total, overflow);
    return hipGetLastError();
}
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // extern "C"
