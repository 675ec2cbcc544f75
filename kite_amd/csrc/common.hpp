// common.hpp -- shared helpers for the gfx950 k-mer engine: error plumbing,
// grow-only device buffers, wave/block scans and the device-wide scan.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace kite {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define KITE_HIP(call)                                                                    \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::kite::Error(std::string(#call) + " failed: " + hipGetErrorString(e_) + \
                                " (" __FILE__ ":" + std::to_string(__LINE__) + ")");     \
    } while (0)

#define KITE_CHECK_LAUNCH() KITE_HIP(hipGetLastError())

// Grow-only device allocation.  Contents are not preserved on growth.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void *ensure(size_t need) {
        if (need == 0) need = 256;
        if (need > bytes) {
            // work queued on the device may still use the old allocation
            if (ptr) KITE_HIP(hipDeviceSynchronize());
            release();
            size_t want = need + need / 8;  // slack so small growth does not reallocate
            KITE_HIP(hipMalloc(&ptr, want));
            bytes = want;
        }
        return ptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ---- device-side scan helpers ---------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Inclusive scan of a 64-bit value across one wave (64 lanes).
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    unsigned lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t o = __shfl_up(v, d, 64);
        if (lane >= (unsigned)d) v += o;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    unsigned lane = __lane_id();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= (unsigned)d) v += o;
    }
    return v;
}

// Exclusive block scan; `ws` is shared scratch of at least THREADS/64 entries.
// Returns the exclusive prefix; *total receives the block sum.
template <int THREADS, class T>
__device__ __forceinline__ T block_excl_scan(T v, T *ws, T *total) {
    constexpr int NW = THREADS / 64;
    unsigned lane = __lane_id();
    unsigned wave = threadIdx.x / 64;
    T inc;
    if constexpr (sizeof(T) == 8) inc = wave_incl_scan_u64(v);
    else inc = wave_incl_scan_u32(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = 0;
        for (int w = 0; w < NW; w++) {
            T t = ws[w];
            ws[w] = s;
            s += t;
        }
        ws[NW] = s;
    }
    __syncthreads();
    T r = ws[wave] + inc - v;
    *total = ws[NW];
    __syncthreads();
    return r;
}

// Device-wide exclusive scan of u32 or u64 counts into u64 offsets; out has
// n + 1 entries (out[n] = total).  Temp holds ceil(n / 4096) + 1 u64.
void scan_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s);
void scan_to_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s);
size_t scan_temp_elems(uint64_t n);

}  // namespace kite
