// scan.hip -- device-wide exclusive scan (reduce, scan of block sums, apply).
// Each block covers 4096 elements: 256 threads x 16 consecutive items.
#include <algorithm>

#include "common.hpp"

namespace kite {

constexpr int SCAN_T = 256;
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = SCAN_T * SCAN_IPT;

size_t scan_temp_elems(uint64_t n) { return ceil_div(n, SCAN_TILE) + 2; }

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const T *__restrict__ in, uint64_t n,
                                                        uint64_t *__restrict__ sums) {
    __shared__ uint64_t ws[SCAN_T / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    // every load unconditional (past n: the last element, not added), so
    // that they are in flight together
    T v[SCAN_IPT];
#pragma unroll
    for (int i = 0; i < SCAN_IPT; i++) {
        const uint64_t idx = base + (uint64_t)i * SCAN_T + threadIdx.x;
        v[i] = in[idx < n ? idx : n - 1];
    }
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; i++)
        if (base + (uint64_t)i * SCAN_T + threadIdx.x < n) acc += (uint64_t)v[i];
    uint64_t total;
    (void)block_excl_scan<SCAN_T, uint64_t>(acc, ws, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums in place; sums[nb] = total.
__global__ __launch_bounds__(1024) void k_scan_sums(uint64_t *__restrict__ sums, uint64_t nb) {
    __shared__ uint64_t ws[1024 / 64 + 1];
    __shared__ uint64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < nb; base += 1024) {
        uint64_t idx = base + threadIdx.x;
        uint64_t v = idx < nb ? sums[idx] : 0;
        uint64_t total;
        uint64_t ex = block_excl_scan<1024, uint64_t>(v, ws, &total);
        if (idx < nb) sums[idx] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const T *__restrict__ in, uint64_t n,
                                                       const uint64_t *__restrict__ sums,
                                                       uint64_t *__restrict__ out) {
    __shared__ uint64_t ws[SCAN_T / 64 + 1];
    uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_IPT;
    uint64_t v[SCAN_IPT];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_IPT; i++) {
        const uint64_t idx = base + i;
        v[i] = (uint64_t)in[idx < n ? idx : n - 1];
    }
#pragma unroll
    for (int i = 0; i < SCAN_IPT; i++) {
        if (base + i >= n) v[i] = 0;
        acc += v[i];
    }
    uint64_t total;
    uint64_t ex = block_excl_scan<SCAN_T, uint64_t>(acc, ws, &total) + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_IPT; i++) {
        uint64_t idx = base + i;
        if (idx < n) out[idx] = ex;
        ex += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = sums[gridDim.x];
}

__global__ void k_zero_one(uint64_t *p) { *p = 0; }

template <class T>
static void scan_impl(const T *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s) {
    if (n == 0) {
        k_zero_one<<<1, 1, 0, s>>>(out);
        KITE_CHECK_LAUNCH();
        return;
    }
    uint64_t nb = ceil_div(n, SCAN_TILE);
    k_scan_reduce<T><<<dim3((unsigned)nb), dim3(SCAN_T), 0, s>>>(in, n, temp);
    KITE_CHECK_LAUNCH();
    k_scan_sums<<<1, 1024, 0, s>>>(temp, nb);
    KITE_CHECK_LAUNCH();
    k_scan_apply<T><<<dim3((unsigned)nb), dim3(SCAN_T), 0, s>>>(in, n, temp, out);
    KITE_CHECK_LAUNCH();
}

// ---- scans whose length is read from device memory -------------------------
// For launch sequences fixed before the data is seen (HIP graphs): the length
// is min(*dn, n_cap), so that the scan stays inside arrays sized for n_cap
// whatever the device word holds; a fixed grid loops over the tiles.

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_reduce_dev(const T *__restrict__ in, const uint64_t *__restrict__ dn,
                                                            uint64_t n_cap, uint64_t *__restrict__ sums) {
    __shared__ uint64_t ws[SCAN_T / 64 + 1];
    const uint64_t n = *dn < n_cap ? *dn : n_cap, nb = ceil_div(n, SCAN_TILE);
    for (uint64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        const uint64_t base = blk * SCAN_TILE;
        T v[SCAN_IPT];
#pragma unroll
        for (int i = 0; i < SCAN_IPT; i++) {
            const uint64_t idx = base + (uint64_t)i * SCAN_T + threadIdx.x;
            v[i] = in[idx < n ? idx : n - 1];
        }
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < SCAN_IPT; i++)
            if (base + (uint64_t)i * SCAN_T + threadIdx.x < n) acc += (uint64_t)v[i];
        uint64_t total;
        (void)block_excl_scan<SCAN_T, uint64_t>(acc, ws, &total);
        if (threadIdx.x == 0) sums[blk] = total;
    }
}

__global__ __launch_bounds__(1024) void k_scan_sums_dev(uint64_t *__restrict__ sums, const uint64_t *__restrict__ dn,
                                                       uint64_t n_cap) {
    __shared__ uint64_t ws[1024 / 64 + 1];
    __shared__ uint64_t carry;
    const uint64_t nb = ceil_div(*dn < n_cap ? *dn : n_cap, SCAN_TILE);
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (uint64_t base = 0; base < nb; base += 1024) {
        const uint64_t idx = base + threadIdx.x;
        const uint64_t v = idx < nb ? sums[idx] : 0;
        uint64_t total;
        const uint64_t ex = block_excl_scan<1024, uint64_t>(v, ws, &total);
        if (idx < nb) sums[idx] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) sums[nb] = carry;
}

template <class T>
__global__ __launch_bounds__(SCAN_T) void k_scan_apply_dev(const T *__restrict__ in, const uint64_t *__restrict__ dn,
                                                           uint64_t n_cap, const uint64_t *__restrict__ sums,
                                                           uint64_t *__restrict__ out) {
    __shared__ uint64_t ws[SCAN_T / 64 + 1];
    const uint64_t n = *dn < n_cap ? *dn : n_cap, nb = ceil_div(n, SCAN_TILE);
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
        return;
    }
    for (uint64_t blk = blockIdx.x; blk < nb; blk += gridDim.x) {
        const uint64_t base = blk * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_IPT;
        uint64_t v[SCAN_IPT];
        uint64_t acc = 0;
#pragma unroll
        for (int i = 0; i < SCAN_IPT; i++) {
            const uint64_t idx = base + i;
            v[i] = (uint64_t)in[idx < n ? idx : n - 1];
        }
#pragma unroll
        for (int i = 0; i < SCAN_IPT; i++) {
            if (base + i >= n) v[i] = 0;
            acc += v[i];
        }
        uint64_t total;
        uint64_t ex = block_excl_scan<SCAN_T, uint64_t>(acc, ws, &total) + sums[blk];
#pragma unroll
        for (int i = 0; i < SCAN_IPT; i++) {
            const uint64_t idx = base + i;
            if (idx < n) out[idx] = ex;
            ex += v[i];
        }
        if (blk == nb - 1 && threadIdx.x == 0) out[n] = sums[nb];
    }
}

template <class T>
static void scan_dev_impl(const T *in, uint64_t *out, const uint64_t *dn, uint64_t n_cap, uint64_t *temp,
                          hipStream_t s) {
    const unsigned g = (unsigned)std::min<uint64_t>(std::max<uint64_t>(1, ceil_div(n_cap, SCAN_TILE)), 2048);
    k_scan_reduce_dev<T><<<dim3(g), dim3(SCAN_T), 0, s>>>(in, dn, n_cap, temp);
    KITE_CHECK_LAUNCH();
    k_scan_sums_dev<<<1, 1024, 0, s>>>(temp, dn, n_cap);
    KITE_CHECK_LAUNCH();
    k_scan_apply_dev<T><<<dim3(g), dim3(SCAN_T), 0, s>>>(in, dn, n_cap, temp, out);
    KITE_CHECK_LAUNCH();
}

void scan_to_u64_dev(const uint32_t *in, uint64_t *out, const uint64_t *dn, uint64_t n_cap, uint64_t *temp,
                     hipStream_t s) {
    scan_dev_impl<uint32_t>(in, out, dn, n_cap, temp, s);
}
void scan_to_u64_dev(const uint64_t *in, uint64_t *out, const uint64_t *dn, uint64_t n_cap, uint64_t *temp,
                     hipStream_t s) {
    scan_dev_impl<uint64_t>(in, out, dn, n_cap, temp, s);
}

void scan_to_u64(const uint32_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s) {
    scan_impl<uint32_t>(in, out, n, temp, s);
}
void scan_to_u64(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t *temp, hipStream_t s) {
    scan_impl<uint64_t>(in, out, n, temp, s);
}

}  // namespace kite
