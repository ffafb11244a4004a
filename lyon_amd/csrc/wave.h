// Wave64 helpers shared by the kernels: lane id, ballots, ds_bpermute shuffles,
// readlane broadcasts, prefix sums and bit tricks.  Every helper must be called from
// wave-uniform control flow.
#pragma once
#include <hip/hip_runtime.h>

#include "lyg_math.h"

// Occupancy target of a kernel, in waves per SIMD (a host-side build of the kernels for
// checking may predefine it empty).
#ifndef LYG_WAVES_PER_EU
#define LYG_WAVES_PER_EU(...) __attribute__((amdgpu_waves_per_eu(__VA_ARGS__)))
#endif

namespace lyg {

static constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int shfl_i(int v, int src) { return __shfl(v, src, WAVE); }
__device__ __forceinline__ uint32_t shfl_u(uint32_t v, int src) {
    return (uint32_t)__shfl((int)v, src, WAVE);
}
__device__ __forceinline__ float shfl_f(float v, int src) { return __shfl(v, src, WAVE); }
__device__ __forceinline__ vec2 shfl_v(vec2 v, int src) {
    return mk(shfl_f(v.x, src), shfl_f(v.y, src));
}
__device__ __forceinline__ uint32_t readlane_u(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float readlane_f(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
// Lane l's value broadcast to the wave (l must be wave-uniform): the result is a scalar,
// so wave-uniform state built from it stays out of vector registers.
__device__ __forceinline__ vec2 readlane_v(vec2 v, int l) {
    return mk(readlane_f(v.x, l), readlane_f(v.y, l));
}
// Inclusive prefix sum over the wave with DPP row moves (no LDS round trips):
// shifts by 1, 2, 4, 8 inside each row of 16 lanes, then row 1 and 3 add the last lane
// of the row before (row_bcast:15) and rows 2-3 add lane 31 (row_bcast:31).  Lanes
// whose source is out of the row read 0 (bound_ctrl); rows outside the row mask add
// the `old` operand, 0.
__device__ __forceinline__ uint32_t incl_scan(uint32_t v) {
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)x;
}
// Minimum of v over the wave, returned to every lane: DPP min-scan (lanes whose source is
// outside the row or the row mask take the `old` operand, 0xffffffff, which cannot win),
// then lane 63's value.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    const int U = -1;
    uint32_t x = v, t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x = x < t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x = x < t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x = x < t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x = x < t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x = x < t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(U, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    x = x < t ? x : t;
    return readlane_u(x, WAVE - 1);
}
// Inclusive running maximum of v over the wave (same DPP pattern as wave_min_u32, with
// 0 as the value that cannot win).
__device__ __forceinline__ uint32_t incl_max_u32(uint32_t v) {
    uint32_t x = v, t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x = x > t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x = x > t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x = x > t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x = x > t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
    x = x > t ? x : t;
    t = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
    x = x > t ? x : t;
    return x;
}
__device__ __forceinline__ int highest_bit(uint64_t m) { return m ? 63 - __builtin_clzll(m) : -1; }
// Last index i in [0, n) with a[i] <= key, for a non-decreasing `a` with a[0] <= key (0
// when n == 0), searched by the whole wave: each round probes 64 evenly spaced entries at
// once, so ceil(log64(n)) dependent loads instead of log2(n).  Wave-uniform arguments.
__device__ __forceinline__ uint32_t wave_search_last_le(const uint32_t* a, uint32_t n,
                                                        uint64_t key) {
    uint32_t lo = 0, hi = n ? n - 1 : 0;
    while (hi > lo) {
        const uint32_t span = hi - lo + 1;
        const uint32_t step = (span + 63u) / 64u;
        const uint32_t idx = lo + (uint32_t)__lane_id() * step;
        const bool ok = idx <= hi && (uint64_t)a[idx] <= key;
        const int k = 63 - __builtin_clzll(__ballot(ok) | 1ull);  // lane 0 probes lo: true
        const uint32_t nlo = lo + (uint32_t)k * step;
        const uint32_t nhi_c = nlo + step - 1u;
        lo = nlo;
        hi = nhi_c < hi ? nhi_c : hi;
        if (step == 1u) break;
    }
    return lo;
}
__device__ __forceinline__ uint64_t mask_lt(int l) { return l >= 64 ? ~0ull : ((1ull << l) - 1ull); }
__device__ __forceinline__ uint64_t mask_le(int l) { return l >= 63 ? ~0ull : ((1ull << (l + 1)) - 1ull); }
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}


}  // namespace lyg
