// lyon_geom flattening as array kernels for gfx950.
//
// Segments (LineSegment / QuadraticBezierSegment / CubicBezierSegment): a count kernel
// evaluates Wang's formula per segment (one lane each), the host-side runtime scans the
// counts, a table kernel walks each segment's serial t sum once and records, for every run
// of 16 consecutive output points, the segment and t of its first point, and an emit
// kernel gives every lane one run: it starts from the table and continues lyon's serial
// sum, evaluating the polynomial form with Horner's rule in lyon's operation order
// (quadratic_bezier.rs:891-938, cubic_bezier.rs:1318-1364).  The points pass through LDS
// so that the stores are fully coalesced.  Without the table (the allocation-free entry
// point) the emit finds the segment of its run in an LDS window of segment offsets and
// recovers t with the closed-form serial chain (lyg_math.h: tchain).
//
// Arcs (Arc::for_each_flattened_with_t, arc.rs:339-368) re-split the remaining arc at
// every step, a serial recurrence, so one lane walks one arc.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lyg_math.h"
#include "lyg_types.h"
#include "stroke_core.h"
#include "arc_core.h"

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

static constexpr uint32_t FL_WAVES = 4;  // waves per block

// Emit without the run table: each wave covers 1,024 consecutive output points; lane l owns
// points [16 l, 16 l + 16) of them.  The wave first loads the offsets of every segment its
// points touch into LDS (each segment has at least one point, so at most 1,025 offsets);
// a lane finds the segment of its first point there, takes t from the closed form (or the
// serial sum near the segment's start) once, then adds the step once per point, which is
// lyon's own serial sum continued.  The points go through LDS so that every store writes
// 64 consecutive points.
static constexpr uint32_t FR_RUN = 16;
static constexpr uint32_t FR_WIN = WAVE * FR_RUN + 1;

// Output points are written once and not read again by the kernel: nontemporal stores
// keep them from displacing the inputs in the caches (measured 1.46x faster than plain
// stores on the emit).
typedef float lyg_f2v __attribute__((ext_vector_type(2)));
typedef float lyg_f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_stream(float2* p, float2 v) {
    const lyg_f2v x = {v.x, v.y};
    __builtin_nontemporal_store(x, reinterpret_cast<lyg_f2v*>(p));
}
__device__ __forceinline__ void store_stream(float4* p, float4 v) {
    const lyg_f4v x = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(x, reinterpret_cast<lyg_f4v*>(p));
}

// Point k of a segment (0-based among its emitted points) at t = the sum of k + 1 steps:
// the polynomial form evaluated as lyon's for_each_flattened_with_t evaluates the curve,
// the last point being the end point itself.
__device__ __forceinline__ vec2 run_point(const SegPoly& sp, uint32_t k, float t) {
    if (k + 1 >= sp.n) return sp.to;
    vec2 v = sp.a0;
    float tt = t;
    v = add(v, mul(sp.a1, tt));
    tt = tt * t;
    v = add(v, mul(sp.a2, tt));
    if (sp.cubic) {
        tt = tt * t;
        v = add(v, mul(sp.a3, tt));
    }
    return v;
}

__global__ void __launch_bounds__(WAVE* FL_WAVES) flatten_emit_runs_kernel(FlattenLaunch L,
                                                                           uint32_t total) {
    __shared__ uint32_t win_all[FL_WAVES][FR_WIN];
    __shared__ float2 stage_all[FL_WAVES][WAVE * (FR_RUN + 1)];  // a run per lane, padded
    const int l = lane_id();
    const uint32_t w = threadIdx.x / WAVE;
    uint32_t* win = win_all[w];
    float2* stage = stage_all[w];
    const uint64_t wave = (uint64_t)blockIdx.x * FL_WAVES + w;
    const uint64_t base = wave * WAVE * FR_RUN;
    if (base >= total) return;  // wave-uniform
    const uint64_t wend = base + WAVE * FR_RUN < total ? base + WAVE * FR_RUN : total;
    // segment of the wave's first point: last segment whose offset <= base
    const uint32_t cur = wave_search_last_le(L.offsets, L.n, base);
    // offsets of segments cur .. until one starts at or beyond the wave's end (L.offsets
    // has n + 1 entries; the last is the total)
    uint32_t nwin = 0;
    for (;;) {
        const uint32_t i = cur + nwin + (uint32_t)l;
        const uint32_t v = i <= L.n ? L.offsets[i] : 0xffffffffu;
        if (nwin + (uint32_t)l < FR_WIN) win[nwin + l] = v;
        const uint32_t last = readlane_u(v, WAVE - 1);
        nwin += WAVE;
        if (last >= wend || nwin >= FR_WIN || cur + nwin > L.n) break;  // uniform
    }
    if (nwin > FR_WIN) nwin = FR_WIN;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const uint64_t g0 = base + (uint64_t)l * FR_RUN;
    if (g0 < wend) {
    const uint32_t g_end = (uint32_t)(g0 + FR_RUN < wend ? g0 + FR_RUN : wend);
    // segment of g0 in the window: last j with win[j] <= g0
    int lo = 0, hi = (int)nwin - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (win[mid] <= g0) lo = mid;
        else hi = mid - 1;
    }
    int j = lo;
    uint32_t seg = cur + (uint32_t)j;
    SegPoly sp = seg_poly(L.kinds[seg], L.ctrl + 8ull * seg, L.tolerance);
    uint32_t seg_begin = win[j], seg_end = win[j + 1];
    // t of the lane's first point (point k + 1 of its segment)
    uint32_t k = (uint32_t)g0 - seg_begin;
    float t = 0.0f;
    if (k + 1 < sp.n) {
        if (k + 1 <= WAVE) {
            for (uint32_t q = 0; q <= k; ++q) t = t + sp.step;
        } else {
            t = tchain(sp.step, k + 1);
        }
    }
    float2* st = stage + (uint32_t)l * (FR_RUN + 1);
    for (uint32_t g = (uint32_t)g0; g < g_end; ++g) {
        if (g >= seg_end) {
            // the next segment (every segment has at least one point)
            ++j;
            seg = cur + (uint32_t)j;
            sp = seg_poly(L.kinds[seg], L.ctrl + 8ull * seg, L.tolerance);
            seg_begin = seg_end;
            seg_end = win[j + 1];
            k = 0;
            t = 0.0f + sp.step;
        }
        const vec2 p = run_point(sp, k, t);
        st[g - (uint32_t)g0] = make_float2(p.x, p.y);
        ++k;
        t = t + sp.step;
    }
    }
    // the wave's points leave LDS in order: 64 consecutive points per store
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    float2* out = reinterpret_cast<float2*>(L.out_points);
#pragma unroll
    for (uint32_t r = 0; r < FR_RUN; ++r) {
        const uint32_t q = r * WAVE + (uint32_t)l;  // point base + q
        if (base + q < wend)
            store_stream(out + base + q, stage[(q / FR_RUN) * (FR_RUN + 1) + q % FR_RUN]);
    }
}

// The run table: for every run of 16 output points, the segment of its first point and
// that point's t by lyon's own serial sum.  One lane walks one segment (t = t + step per
// point, the point k of a segment taking the sum of k + 1 steps) and writes the entry of
// every point whose index is a multiple of 16; a segment's last point takes its end point,
// so its t is never read.  The same pass stores every segment's polynomial form, step and
// count (48 bytes), so the emit loads them instead of evaluating Wang's formula again.
__device__ __forceinline__ void poly_store(float4* dst, const SegPoly& sp) {
    dst[0] = make_float4(sp.a0.x, sp.a0.y, sp.a1.x, sp.a1.y);
    dst[1] = make_float4(sp.a2.x, sp.a2.y, sp.a3.x, sp.a3.y);
    dst[2] = make_float4(sp.to.x, sp.to.y, sp.step, u2f(sp.n | (sp.cubic ? 0x80000000u : 0u)));
}
__device__ __forceinline__ SegPoly poly_load(const float4* src) {
    const float4 a = src[0], b = src[1], c = src[2];
    SegPoly sp;
    sp.a0 = mk(a.x, a.y);
    sp.a1 = mk(a.z, a.w);
    sp.a2 = mk(b.x, b.y);
    sp.a3 = mk(b.z, b.w);
    sp.to = mk(c.x, c.y);
    sp.step = c.z;
    const uint32_t nc = f2u(c.w);
    sp.n = nc & 0x7fffffffu;
    sp.cubic = (nc >> 31) != 0;
    return sp;
}

__global__ void flatten_runs_table_kernel(FlattenLaunch L) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const SegPoly sp = seg_poly(L.kinds[i], L.ctrl + 8ull * i, L.tolerance);
    poly_store(reinterpret_cast<float4*>(L.polys) + 3ull * i, sp);
    const uint32_t off = L.offsets[i];
    uint2* runs = reinterpret_cast<uint2*>(L.runs);
    // k: the segment's first point that starts a run; t: the sum of k + 1 steps
    uint32_t k = (FR_RUN - (off & (FR_RUN - 1))) & (FR_RUN - 1);
    float t = 0.0f;
    for (uint32_t q = 0; q <= k; ++q) t = t + sp.step;
    for (; k < sp.n; k += FR_RUN) {
        runs[(off + k) / FR_RUN] = make_uint2(i, f2u(t));
#pragma unroll
        for (uint32_t q = 0; q < FR_RUN; ++q) t = t + sp.step;
    }
}

// Emit from the run table: lane l of a wave owns points [16 l, 16 l + 16) of the wave's
// 1,024; it reads its run's segment and t from the table, the segment's polynomial form
// and its two offsets, and continues the serial sum with one addition per point (moving
// to the next segment, whose offset it reads, when the run crosses one).  The points go
// through LDS so that every store writes 128 consecutive points (64 with V16 false, for an
// output that is only 8-byte aligned).
template <bool V16>
__global__ void __launch_bounds__(WAVE* FL_WAVES) flatten_emit_table_kernel(FlattenLaunch L,
                                                                            uint32_t total) {
    constexpr uint32_t PAD = V16 ? 2 : 1;  // V16: 16-byte aligned pairs in LDS
    __shared__ __align__(16) float2 stage_all[FL_WAVES][WAVE * (FR_RUN + PAD)];  // a run per lane, padded
    const int l = lane_id();
    const uint32_t w = threadIdx.x / WAVE;
    float2* stage = stage_all[w];
    const uint64_t wave = (uint64_t)blockIdx.x * FL_WAVES + w;
    const uint64_t base = wave * WAVE * FR_RUN;
    if (base >= total) return;  // wave-uniform
    const uint64_t wend = base + WAVE * FR_RUN < total ? base + WAVE * FR_RUN : total;
    const uint64_t g0 = base + (uint64_t)l * FR_RUN;
    if (g0 < wend) {
        const uint32_t g_end = (uint32_t)(g0 + FR_RUN < wend ? g0 + FR_RUN : wend);
        const uint2 run = reinterpret_cast<const uint2*>(L.runs)[g0 / FR_RUN];
        uint32_t seg = run.x;
        float t = u2f(run.y);
        const float4* polys = reinterpret_cast<const float4*>(L.polys);
        SegPoly sp = poly_load(polys + 3ull * seg);
        uint32_t k = (uint32_t)g0 - L.offsets[seg];
        uint32_t seg_end = L.offsets[seg + 1];
        float2* st = stage + (uint32_t)l * (FR_RUN + PAD);
        for (uint32_t g = (uint32_t)g0; g < g_end; ++g) {
            if (g >= seg_end) {
                // the next segment (every segment has at least one point)
                ++seg;
                sp = poly_load(polys + 3ull * seg);
                seg_end = L.offsets[seg + 1];
                k = 0;
                t = 0.0f + sp.step;
            }
            const vec2 p = run_point(sp, k, t);
            st[g - (uint32_t)g0] = make_float2(p.x, p.y);
            ++k;
            t = t + sp.step;
        }
    }
    // the wave's points leave LDS in order: 64 consecutive points per store
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    float2* out = reinterpret_cast<float2*>(L.out_points);
    if constexpr (!V16) {
#pragma unroll
        for (uint32_t r = 0; r < FR_RUN; ++r) {
            const uint32_t q = r * WAVE + (uint32_t)l;  // point base + q
            if (base + q < wend)
                store_stream(out + base + q, stage[(q / FR_RUN) * (FR_RUN + PAD) + q % FR_RUN]);
        }
    } else {
        // two points per lane and store: 128 consecutive points per store
        const uint32_t nw = (uint32_t)(wend - base);
#pragma unroll
        for (uint32_t r = 0; r < FR_RUN / 2; ++r) {
            const uint32_t q = 2 * (r * WAVE + (uint32_t)l);  // points base + q, q + 1
            const float4 v = *reinterpret_cast<const float4*>(
                stage + (q / FR_RUN) * (FR_RUN + PAD) + q % FR_RUN);
            if (q + 1 < nw) store_stream(reinterpret_cast<float4*>(out + base + q), v);
            else if (q < nw) store_stream(out + base + q, make_float2(v.x, v.y));
        }
    }
}

// ---- arcs -------------------------------------------------------------------------------
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
    const float* a = arcs + 7ull * i;
    ArcD arc{mk(a[0], a[1]), mk(a[2], a[3]), a[4], a[5], a[6]};
    if (counts) {
        counts[i] = (uint32_t)arc_quad_count(arc);
        return;
    }
    float* o = out + 6ull * offsets[i];
    arc_to_quads_each(arc, [&](int k, vec2 from, vec2 ctrl, vec2 to) {
        o[6 * k + 0] = from.x;
        o[6 * k + 1] = from.y;
        o[6 * k + 2] = ctrl.x;
        o[6 * k + 3] = ctrl.y;
        o[6 * k + 4] = to.x;
        o[6 * k + 5] = to.y;
    });
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

// Large scans (n > SCAN_SMALL, lyg_launch_scan_tiles): three launches.  Tiles of SCAN_TILE counts (256 threads x
// 16 consecutive counts each): per-tile sums, one block scanning the tile sums, then every
// tile scanned again with its prefix and written.
static constexpr uint32_t SCAN_SMALL = 32768;
static constexpr uint32_t SCAN_TILE = 4096;

__device__ __forceinline__ uint64_t block_incl_scan_u64(uint64_t v, uint64_t* sh) {
    // 256 threads: wave-level inclusive scan, then the wave totals
    const int l = lane_id();
    const int w = threadIdx.x / WAVE;
    for (int o = 1; o < WAVE; o <<= 1) {
        const uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
        const int src = l >= o ? l - o : l;
        const uint64_t t = (uint64_t)shfl_u(lo, src) | ((uint64_t)shfl_u(hi, src) << 32);
        if (l >= o) v += t;
    }
    if (l == WAVE - 1) sh[w] = v;
    __syncthreads();
    uint64_t before = 0;
    for (int k = 0; k < w; ++k) before += sh[k];
    __syncthreads();
    return v + before;
}

__global__ void __launch_bounds__(256) scan_tile_sums_kernel(const uint32_t* counts, uint32_t n,
                                                             uint64_t* tile_sums) {
    __shared__ uint64_t sh[4];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint64_t v = 0;
    for (uint32_t i = 0; i < SCAN_TILE / 256; ++i) {
        const uint64_t k = base + (uint64_t)i * 256 + threadIdx.x;
        if (k < n) v += counts[k];
    }
    v = block_incl_scan_u64(v, sh);
    if (threadIdx.x == 255) tile_sums[blockIdx.x] = v;
}

__global__ void __launch_bounds__(256) scan_tile_prefix_kernel(uint64_t* tile_sums, uint32_t ntiles,
                                                               uint64_t* total) {
    __shared__ uint64_t sh[4];
    uint64_t carry = 0;
    for (uint32_t b = 0; b < ntiles; b += 256) {
        const uint32_t k = b + threadIdx.x;
        const uint64_t v = k < ntiles ? tile_sums[k] : 0;
        const uint64_t incl = block_incl_scan_u64(v, sh);
        if (k < ntiles) tile_sums[k] = carry + incl - v;  // exclusive prefix
        __syncthreads();
        if (threadIdx.x == 255) sh[0] = incl;
        __syncthreads();
        carry += sh[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(256) scan_tile_write_kernel(const uint32_t* counts, uint32_t n,
                                                              const uint64_t* tile_prefix,
                                                              uint32_t* offsets, uint32_t* overflow) {
    __shared__ uint64_t sh[4];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * 16;
    uint32_t c[16];
    uint64_t sum = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        c[i] = base + i < n ? counts[base + i] : 0u;
        sum += c[i];
    }
    const uint64_t incl = block_incl_scan_u64(sum, sh);
    uint64_t run = tile_prefix[blockIdx.x] + incl - sum;
    bool ovf = false;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        if (base + i < n) {
            if (run > 0xffffffffull) ovf = true;
            offsets[base + i] = (uint32_t)run;
        }
        run += c[i];
    }
    if (ovf) atomicOr(overflow, 1u);
}

__global__ void scan_finish_kernel(const uint64_t* total, uint32_t* last, uint32_t* overflow) {
    if (threadIdx.x != 0) return;
    const uint64_t t = *total;
    *last = (uint32_t)t;
    if (t > 0xffffffffull) atomicOr(overflow, 1u);
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
// With L->runs: the run table first (8 bytes per 16 points, 48 per segment), then the
// emit starting its runs from it; without: the emit recovers t in closed form.
hipError_t lyg_launch_flatten_emit(const lyg::FlattenLaunch* L, uint32_t total, hipStream_t s) {
    if (total == 0) return hipSuccess;
    const uint64_t per_block = (uint64_t)lyg::WAVE * lyg::FR_RUN * lyg::FL_WAVES;
    const dim3 grid((uint32_t)((total + per_block - 1) / per_block));
    if (L->runs) {
        hipLaunchKernelGGL(lyg::flatten_runs_table_kernel, dim3((L->n + 255) / 256), dim3(256),
                           0, s, *L);
        // 16-byte stores need a 16-byte aligned output (every device allocation is)
        if ((uintptr_t)L->out_points & 15)
            hipLaunchKernelGGL(lyg::flatten_emit_table_kernel<false>, grid,
                               dim3(lyg::WAVE * lyg::FL_WAVES), 0, s, *L, total);
        else
            hipLaunchKernelGGL(lyg::flatten_emit_table_kernel<true>, grid,
                               dim3(lyg::WAVE * lyg::FL_WAVES), 0, s, *L, total);
    } else {
        hipLaunchKernelGGL(lyg::flatten_emit_runs_kernel, grid,
                           dim3(lyg::WAVE * lyg::FL_WAVES), 0, s, *L, total);
    }
    return hipGetLastError();
}
size_t lyg_flatten_runs_bytes(uint32_t total) {
    return 8ull * ((uint64_t)total / lyg::FR_RUN + 1);
}
size_t lyg_flatten_poly_bytes(uint32_t n) { return 48ull * n; }
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
// Bytes of the `tiles` scratch lyg_launch_scan_tiles needs for n counts.
size_t lyg_scan_tiles_bytes(uint32_t n) {
    return 8ull * ((n + lyg::SCAN_TILE - 1) / lyg::SCAN_TILE + 1);
}
// The same scan for large n, across the whole GPU (tiles: lyg_scan_tiles_bytes(n) bytes).
hipError_t lyg_launch_scan_tiles(const uint32_t* counts, uint32_t n, uint32_t* offsets,
                                 uint64_t* total, uint32_t* overflow, uint64_t* tiles,
                                 hipStream_t s) {
    if (n <= lyg::SCAN_SMALL) return lyg_launch_scan(counts, n, offsets, total, overflow, s);
    const uint32_t ntiles = (n + lyg::SCAN_TILE - 1) / lyg::SCAN_TILE;
    hipLaunchKernelGGL(lyg::scan_tile_sums_kernel, dim3(ntiles), dim3(256), 0, s, counts, n, tiles);
    hipLaunchKernelGGL(lyg::scan_tile_prefix_kernel, dim3(1), dim3(256), 0, s, tiles, ntiles, total);
    hipLaunchKernelGGL(lyg::scan_tile_write_kernel, dim3(ntiles), dim3(256), 0, s, counts, n, tiles,
                       offsets, overflow);
    hipLaunchKernelGGL(lyg::scan_finish_kernel, dim3(1), dim3(64), 0, s, total, offsets + n, overflow);
    return hipGetLastError();
}
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(lyg::fill_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

}  // extern "C"

// ---- fixed scratch of the allocation-free flatten entry point -------------------------
// lyg_flatten_segments_device_noalloc keeps its per-segment counts, the tile sums of its
// offset scan, its totals and its overflow flag in this array instead of buffers it would
// allocate.  Every GPU holds its own copy for the life of the process; host code reaches
// it only through the address hipGetSymbolAddress returns for the current device (the
// array's name in host code is not a device address).
namespace {
constexpr uint32_t NOALLOC_MAX_SEGMENTS = 1u << 20;
constexpr size_t NOALLOC_COUNTERS_AT = 0;    // u32[16]: [0] scan overflow flag
constexpr size_t NOALLOC_TOTALS_AT = 256;    // u64[4]: [0] flattened points
constexpr size_t NOALLOC_COUNTS_AT = 512;    // u32[NOALLOC_MAX_SEGMENTS]
constexpr size_t NOALLOC_TILES_AT = NOALLOC_COUNTS_AT + 4ull * NOALLOC_MAX_SEGMENTS;
constexpr size_t NOALLOC_TILES_BYTES =
    8ull * ((NOALLOC_MAX_SEGMENTS + lyg::SCAN_TILE - 1) / lyg::SCAN_TILE + 1);
constexpr size_t NOALLOC_SCRATCH_BYTES = (NOALLOC_TILES_AT + NOALLOC_TILES_BYTES + 255) & ~size_t(255);
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
void* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

extern "C" {

uint32_t lyg_noalloc_max_segments() { return NOALLOC_MAX_SEGMENTS; }

// Once per device, with that device current (lyg_ctx_create calls it): looks up this
// device's copy of `scratch` and keeps its address.
hipError_t lyg_noalloc_scratch_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (noalloc_ptr[device]) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    noalloc_ptr[device] = p;
    return hipSuccess;
}

// The scratch layout for `device` (null pointers before lyg_noalloc_scratch_init).
void lyg_noalloc_scratch(int device, uint32_t** counters, uint64_t** totals, uint32_t** counts,
                         uint64_t** tiles) {
    uint8_t* base = nullptr;
    if (device >= 0 && device < NOALLOC_MAX_DEVICES) {
        std::lock_guard<std::mutex> g(noalloc_mu);
        base = static_cast<uint8_t*>(noalloc_ptr[device]);
    }
    *counters = base ? reinterpret_cast<uint32_t*>(base + NOALLOC_COUNTERS_AT) : nullptr;
    *totals = base ? reinterpret_cast<uint64_t*>(base + NOALLOC_TOTALS_AT) : nullptr;
    *counts = base ? reinterpret_cast<uint32_t*>(base + NOALLOC_COUNTS_AT) : nullptr;
    *tiles = base ? reinterpret_cast<uint64_t*>(base + NOALLOC_TILES_AT) : nullptr;
}

}  // extern "C"
