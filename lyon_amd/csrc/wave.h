// Wave64 helpers shared by the kernels: lane id, ballots, ds_bpermute shuffles,
// readlane broadcasts, prefix sums and bit tricks.  Every helper must be called from
// wave-uniform control flow.
#pragma once
#include <hip/hip_runtime.h>

#include "lyg_math.h"

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
__device__ __forceinline__ uint32_t incl_scan(uint32_t v) {
    int l = lane_id();
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t t = shfl_u(v, l >= o ? l - o : l);
        if (l >= o) v += t;
    }
    return v;
}
__device__ __forceinline__ int highest_bit(uint64_t m) { return m ? 63 - __builtin_clzll(m) : -1; }
__device__ __forceinline__ uint64_t mask_lt(int l) { return l >= 64 ? ~0ull : ((1ull << l) - 1ull); }
__device__ __forceinline__ uint64_t mask_le(int l) { return l >= 63 ? ~0ull : ((1ull << (l + 1)) - 1ull); }
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}


}  // namespace lyg
