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

// Device memory a grown buffer gave up.  Freeing device memory makes the
// driver wipe it on the same copy hardware the process's device-to-host
// copies use: they run at about half rate while it does (about 30 ms per
// GiB freed).  Allocations of RETIRE_MIN bytes or more that a growing
// DevBuf replaces are therefore kept, still allocated, and freed together
// by free_retired: at a context's close, when an allocation fails for
// lack of memory (then retried), or at once past RETIRE_CAP bytes kept.
// Nothing waits for the device when a buffer is retired: queued work may
// still read it, and free_retired waits for the device first.
constexpr size_t RETIRE_MIN = 64ull << 20;
constexpr size_t RETIRE_CAP = 128ull << 30;
void retire_device(void *p, size_t bytes);
void free_retired();

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
            if (ptr && bytes >= RETIRE_MIN) {
                retire_device(ptr, bytes);
                ptr = nullptr;
                bytes = 0;
            }
            // work queued on the device may still use the old allocation
            if (ptr) KITE_HIP(hipDeviceSynchronize());
            release();
            size_t want = need + need / 8;  // slack so small growth does not reallocate
            hipError_t e = hipMalloc(&ptr, want);
            if (e == hipErrorOutOfMemory) {
                // retired buffers hold memory that no queued work needs
                (void)hipGetLastError();
                free_retired();
                e = hipMalloc(&ptr, want);
            }
            if (e != hipSuccess) {
                ptr = nullptr;
                throw Error("hipMalloc of " + std::to_string(want) + " bytes failed: " + hipGetErrorString(e));
            }
            bytes = want;
        }
        return ptr;
    }
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ---- device-side scan helpers ---------------------------------------------

__device__ __forceinline__ uint64_t lanemask_lt() {
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Wave scans by DPP lane moves (no LDS round trips): row_shr:1/2/4/8 scan
// each row of 16 lanes, then row_bcast:15 adds the end of row 0 (2) to row
// 1 (3), and row_bcast:31 adds the end of rows 0-1 to rows 2-3.  Lanes whose
// source is outside the row keep 0.  Every lane of the wave must be active.
template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xF, false);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const uint32_t lo = dpp_u32<CTRL, ROWS>((uint32_t)v);
    const uint32_t hi = dpp_u32<CTRL, ROWS>((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Inclusive scan of a 64-bit value across one wave (64 lanes).
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    v += dpp_u64<0x111, 0xF>(v);
    v += dpp_u64<0x112, 0xF>(v);
    v += dpp_u64<0x114, 0xF>(v);
    v += dpp_u64<0x118, 0xF>(v);
    v += dpp_u64<0x142, 0xA>(v);
    v += dpp_u64<0x143, 0xC>(v);
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
    v += dpp_u32<0x111, 0xF>(v);
    v += dpp_u32<0x112, 0xF>(v);
    v += dpp_u32<0x114, 0xF>(v);
    v += dpp_u32<0x118, 0xF>(v);
    v += dpp_u32<0x142, 0xA>(v);
    v += dpp_u32<0x143, 0xC>(v);
    return v;
}

// Exclusive block scan; `ws` is shared scratch of at least THREADS/64 + 1
// entries.  Returns the exclusive prefix; *total receives the block sum.
// Wave 0 scans the wave totals.  With `reuse` false the closing barrier is
// left out: the caller then passes another barrier before `ws` is written
// again.
template <int THREADS, class T, bool reuse = true>
__device__ __forceinline__ T block_excl_scan(T v, T *ws, T *total) {
    constexpr int NW = THREADS / 64;
    static_assert(NW <= 64, "one wave scans the wave totals");
    const unsigned lane = threadIdx.x & 63;
    const unsigned wave = threadIdx.x / 64;
    T inc;
    if constexpr (sizeof(T) == 8) inc = wave_incl_scan_u64(v);
    else inc = wave_incl_scan_u32(v);
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    if (wave == 0) {
        const T t = lane < (unsigned)NW ? ws[lane] : T(0);
        T s;
        if constexpr (sizeof(T) == 8) s = wave_incl_scan_u64(t);
        else s = wave_incl_scan_u32(t);
        if (lane < (unsigned)NW) ws[lane] = s - t;
        if (lane == (unsigned)NW - 1) ws[NW] = s;
    }
    __syncthreads();
    T r = ws[wave] + inc - v;
    *total = ws[NW];
    if constexpr (reuse) __syncthreads();
    return r;
}

// Device-wide exclusive scan of u32 or u64 counts into u64 offsets; out has
// n + 1 entries (out[n] = total).  Temp holds ceil(n / 4096) + 1 u64.
void scan_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s);
void scan_to_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s);
size_t scan_temp_elems(uint64_t n);
// The same with the length read from device memory (min(*dn, n_cap)), for
// launch sequences fixed before the data is seen; temp as for n_cap.
void scan_to_u64_dev(const uint32_t *in, uint64_t *out, const uint64_t *dn, uint64_t n_cap, uint64_t *temp,
                     hipStream_t s);
void scan_to_u64_dev(const uint64_t *in, uint64_t *out, const uint64_t *dn, uint64_t n_cap, uint64_t *temp,
                     hipStream_t s);

}  // namespace kite
