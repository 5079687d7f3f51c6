// Shared host/device helpers for the FyzzXYZZY HIP library (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace fyzz_hip {

using u32 = std::uint32_t;
using u64 = std::uint64_t;

// Status codes returned across the C-ABI.
enum Status : int {
    kOk = 0,
    kInvalidArgument = -1,
    kDeviceError = -2,
    kOverflow = -3,
    kLogicError = -4,
    kOutOfMemory = -5,
};

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define FYZZ_HIP_CHECK(expr)                                                          \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::fyzz_hip::Error(e_ == hipErrorOutOfMemory ? ::fyzz_hip::kOutOfMemory \
                                                              : ::fyzz_hip::kDeviceError, \
                                    std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define FYZZ_REQUIRE(cond, code, msg)                     \
    do {                                                  \
        if (!(cond)) throw ::fyzz_hip::Error((code), (msg)); \
    } while (0)

// Owning device allocation that only grows.
template <class T>
struct DevBuf {
    T* p = nullptr;
    std::size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void reserve(std::size_t n) {
        if (n <= cap) return;
        if (p) FYZZ_HIP_CHECK(hipFree(p));
        p = nullptr;
        cap = 0;
        FYZZ_HIP_CHECK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
        cap = n;
    }
    T* get() const { return p; }
};

constexpr int kWave = 64;

inline unsigned grid_for(std::size_t n, unsigned block) {
    return static_cast<unsigned>((n + block - 1) / block);
}

#if defined(__HIP_DEVICE_COMPILE__) || defined(__HIPCC__)
// popcount(mask & lanes-below-me): the wave64 rank of this lane in `mask`.
__device__ inline u32 lane_rank(u64 mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<u32>(mask), 0u));
}
__device__ inline u32 lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Inclusive wave64 prefix sum.
__device__ inline u32 wave_inclusive_scan(u32 v) {
    const u32 lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const u32 o = __shfl_up(v, d, kWave);
        if (lane >= static_cast<u32>(d)) v += o;
    }
    return v;
}
__device__ inline u64 wave_inclusive_scan64(u64 v) {
    const u32 lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const u64 o = __shfl_up(v, d, kWave);
        if (lane >= static_cast<u32>(d)) v += o;
    }
    return v;
}

// Exclusive scan over a block of BLOCK threads (multiple of 64). `scratch`
// needs BLOCK/64 + 1 words of LDS. Returns the exclusive prefix and writes
// the block total. Contains __syncthreads.
template <int BLOCK>
__device__ inline u32 block_exclusive_scan(u32 v, u32* scratch, u32* total) {
    constexpr int W = BLOCK / kWave;
    const u32 lane = lane_id();
    const u32 wave = threadIdx.x / kWave;
    const u32 inc = wave_inclusive_scan(v);
    if (lane == kWave - 1) scratch[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 run = 0;
        for (int w = 0; w < W; ++w) {
            const u32 t = scratch[w];
            scratch[w] = run;
            run += t;
        }
        scratch[W] = run;
    }
    __syncthreads();
    const u32 r = scratch[wave] + inc - v;
    *total = scratch[W];
    __syncthreads();
    return r;
}
#endif

}  // namespace fyzz_hip
