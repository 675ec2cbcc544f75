// corr.hip -- Pearson (log1p counts) and Spearman (average-tie ranks)
// correlation of union count columns, and the FP64 Jacobi PCA.
//
// Reference semantics (src/analyze/mod.rs): values are ln(1 + count)
// (:780-784); Pearson r = centered cross / sqrt(centered sumsq products),
// 0 when the denominator is <= 0 (:222-236); Spearman = Pearson of 1-based
// average-tie ranks over all union rows, zeros included (:406-438,
// :559-578); covariance divides by N-1 (:238-252); PCA scores are
// v * sqrt(max(lambda, 0)) with eigenvalues descending (:1605-1625).
//
// Every value depends only on (column, count), so the column statistics
// come exactly from per-column count histograms: means and variances in
// FP64, ranks as cum(<c) + (h(c) + 1) / 2.  The only per-row work is the
// centered Gram G = Z^T Z of standardized values z = (v - mean) / sd, done
// with bf16 MFMA on a hi/lo split of z (hi*hi + hi*lo + lo*hi, the last two
// as one product whose transpose is added at the end; FP32 inside a K-tile,
// FP64 across tiles), or with FP64 MFMA as the baseline.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "corr.hpp"

namespace kite {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

constexpr int HIST_SMALL = 256;   // LDS-aggregated count values per column
constexpr int MAX_COLS = 32;      // one 32x32 MFMA tile of columns
constexpr int GT = 256;           // Gram workgroup threads (4 waves)
constexpr int FLUSH_STEPS = 32;   // bf16 steps (16 rows each) per FP32 flush: 512 rows

// ---- rank codes ------------------------------------------------------------
// A cell's code is the rank of its count among the distinct counts of its
// column.  Counts below `small` are ranked from a per-column histogram;
// the (rare) larger ones from a sorted per-column list built on the host.

// Largest cell of a u32 or u64 count matrix.
template <class T>
__global__ void k_max_cells(const T *__restrict__ p, uint64_t n, unsigned long long *__restrict__ out) {
    uint64_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = m > (uint64_t)p[i] ? m : (uint64_t)p[i];
    for (int d = 32; d > 0; d >>= 1) {
        const uint64_t o = __shfl_xor(m, d, 64);
        m = m > o ? m : o;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)m);
}

// Per-column histogram of the counts below `small` (blockIdx.y: a group of
// MAX_COLS columns; counts below HIST_SMALL aggregated in LDS first), and
// the number of cells at or above `small`.
template <class T>
__global__ __launch_bounds__(256) void k_rank_hist(const T *__restrict__ cells, uint64_t N, uint32_t n, uint32_t small,
                                                   uint32_t *__restrict__ hist, unsigned long long *__restrict__ n_big) {
    __shared__ uint32_t sm[MAX_COLS * HIST_SMALL];
    __shared__ uint32_t sbig;
    const uint32_t c0 = blockIdx.y * MAX_COLS;
    const uint32_t nc = min((uint32_t)MAX_COLS, n - c0);
    for (uint32_t i = threadIdx.x; i < (uint32_t)(MAX_COLS * HIST_SMALL); i += 256) sm[i] = 0;
    if (threadIdx.x == 0) sbig = 0;
    __syncthreads();
    const uint64_t total = N * (uint64_t)nc;
    uint32_t big = 0;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        const uint64_t r = e / nc;
        const uint32_t j = (uint32_t)(e - r * nc);
        const uint64_t v = (uint64_t)cells[r * n + c0 + j];
        if (v < (uint64_t)HIST_SMALL && v < (uint64_t)small) atomicAdd(&sm[j * HIST_SMALL + (uint32_t)v], 1u);
        else if (v < (uint64_t)small) atomicAdd(&hist[(uint64_t)(c0 + j) * small + v], 1u);
        else big++;
    }
    if (big) atomicAdd(&sbig, big);
    __syncthreads();
    const uint32_t lim = min((uint32_t)HIST_SMALL, small);
    for (uint32_t i = threadIdx.x; i < nc * HIST_SMALL; i += 256) {
        const uint32_t j = i / HIST_SMALL, v = i % HIST_SMALL;
        if (v < lim && sm[i]) atomicAdd(&hist[(uint64_t)(c0 + j) * small + v], sm[i]);
    }
    if (threadIdx.x == 0 && sbig) atomicAdd(n_big, (unsigned long long)sbig);
}

// The cells at or above `small` as (column, count) pairs, in any order.
struct BigCell {
    uint64_t col, value;
};

template <class T>
__global__ void k_rank_big(const T *__restrict__ cells, uint64_t cells_n, uint32_t n, uint32_t small,
                           BigCell *__restrict__ out, uint64_t cap, unsigned long long *__restrict__ cursor) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells_n) return;
    const uint64_t v = (uint64_t)cells[i];
    if (v < (uint64_t)small) return;
    const uint64_t at = atomicAdd(cursor, 1ull);
    if (at < cap) out[at] = BigCell{i % n, v};
}

// Rank codes: map[column][count] for counts below `small`; above, the
// column's count of small distinct values plus the position in its sorted
// list of large distinct values.
template <class T, class C>
__global__ void k_rank_codes(const T *__restrict__ cells, uint64_t cells_n, uint32_t n, uint32_t small,
                             const uint32_t *__restrict__ map, const uint64_t *__restrict__ big_off,
                             const uint64_t *__restrict__ big_vals, const uint32_t *__restrict__ big_base,
                             C *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells_n) return;
    const uint32_t j = (uint32_t)(i % n);
    const uint64_t v = (uint64_t)cells[i];
    uint32_t code;
    if (v < (uint64_t)small) {
        code = map[(uint64_t)j * small + v];
    } else {
        uint64_t lo = big_off[j], hi = big_off[j + 1];
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            if (big_vals[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        code = big_base[j] + (uint32_t)(lo - big_off[j]);
    }
    out[i] = (C)code;
}

// A 32-column code matrix of up to two 16-column blocks of the n-column
// one: out[r][c] = codes[r][cols[c]], or code 0 where cols[c] < 0 (a
// padding column, whose value table is all zero).
template <class C>
__global__ void k_gather_cols(const C *__restrict__ codes, uint64_t N, uint32_t n, const int *__restrict__ cols,
                              C *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N * MAX_COLS) return;
    const uint64_t r = i / MAX_COLS;
    const int c = cols[i % MAX_COLS];
    out[i] = c >= 0 ? codes[r * n + (uint32_t)c] : (C)0;
}

__device__ __forceinline__ __bf16 bf16_from_bits(uint16_t b) {
    __bf16 r;
    memcpy(&r, &b, 2);
    return r;
}

// ---- Gram kernels --------------------------------------------------------
// Each wave owns a contiguous range of rows and walks it in blocks of RB
// rows.  The codes of the next block are loaded into registers (coalesced
// 32-bit words) while the current block is multiplied; a block passes
// through the wave's own LDS area so that every lane can read the (row,
// column) cells its MFMA operands need.  The value table of the first LSM
// counts of every column lives in LDS; larger counts read the global table.
constexpr int RB = 64;    // rows per wave block
constexpr int LSM = 256;  // table entries per column held in LDS (<= stride)
constexpr int LSM64 = 128;  // the same for the FP64 table (8-byte entries)
// Column strides of the LDS tables, one entry longer than the table: lanes
// that read different columns at the same (mostly small) count then fall
// into different LDS banks.
constexpr int LSP = LSM + 1;

template <class T>
struct CodeBlock {
    static constexpr int WORDS = RB * MAX_COLS * (int)sizeof(T) / 4;  // LDS words per wave
    static constexpr int MAXW = WORDS / 64;                          // words per lane
};

// Loads the cnt codes starting at element e0 (e0 * sizeof(T) is 4-byte
// aligned) as 32-bit words; words past the block read as zero.
template <class T>
__device__ __forceinline__ void fetch_codes(const T *__restrict__ codes, uint64_t e0, uint32_t cnt,
                                            uint32_t (&w)[CodeBlock<T>::MAXW], int lane) {
    const uint32_t nbytes = cnt * (uint32_t)sizeof(T);
    const uint32_t nfull = nbytes / 4;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(codes + e0);
#pragma unroll
    for (int j = 0; j < CodeBlock<T>::MAXW; j++) {
        const uint32_t idx = (uint32_t)j * 64u + (uint32_t)lane;
        uint32_t v = 0;
        if (idx < nfull) {
            v = src[idx];
        } else if (idx == nfull && (nbytes & 3u)) {  // trailing partial word: element by element
            const uint32_t first = nfull * 4u / (uint32_t)sizeof(T);
            for (uint32_t q = first; q < cnt; q++)
                v |= (uint32_t)codes[e0 + q] << (8u * (uint32_t)sizeof(T) * (q - first));
        }
        w[j] = v;
    }
}

template <class T>
__device__ __forceinline__ void stage_codes(uint32_t *__restrict__ sw, const uint32_t (&w)[CodeBlock<T>::MAXW],
                                            int lane) {
#pragma unroll
    for (int j = 0; j < CodeBlock<T>::MAXW; j++) sw[j * 64 + lane] = w[j];
}

template <class T>
__device__ __forceinline__ uint32_t staged_code(const uint32_t *__restrict__ sw, uint32_t e) {
    if constexpr (sizeof(T) == 1) return reinterpret_cast<const uint8_t *>(sw)[e];
    else if constexpr (sizeof(T) == 2) return reinterpret_cast<const uint16_t *>(sw)[e];
    else return sw[e];
}

// lut: `stride` entries per column, each holding the NM methods' values as
// packed bf16 bits (hi | lo2 << 16), where z = hi + lo2 / 2 + O(2^-16 z):
// lo2 is twice the low part, which is exact in bf16.  The MFMAs accumulate
// A = H^T H + H^T Lo2 (+ Lo^T Lo), so that (A + A^T) / 2 = H^T H + H^T L +
// L^T H (+ L^T L) is the split Gram with two products per step instead of
// three; the host symmetrizes.  partial: [wave][NM][16][64] FP64
// accumulators in MFMA output layout.
template <int NM> struct GramEntry { using type = uint32_t; };
template <> struct GramEntry<2> { using type = uint2; };

template <class T, int NM>
__global__ __launch_bounds__(GT, 2) void k_gram_bf16(const T *__restrict__ codes, uint64_t N, int n,
                                                   const uint32_t *__restrict__ lut, uint32_t stride,
                                                   uint64_t rows_per_wave, int with_lolo,
                                                   double *__restrict__ partial) {
    using Entry = typename GramEntry<NM>::type;
    constexpr uint32_t NONE = 0xFFFFFFFFu;
    const Entry *glut = reinterpret_cast<const Entry *>(lut);
    __shared__ Entry slut[MAX_COLS * LSP];
    __shared__ uint32_t sblk[GT / 64][CodeBlock<T>::WORDS];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < n * LSM; i += GT)
        slut[(i / LSM) * LSP + (i % LSM)] = glut[(uint64_t)(i / LSM) * stride + (i % LSM)];
    __syncthreads();
    const uint64_t wave = (uint64_t)blockIdx.x * (GT / 64) + wv;
    const uint64_t r0 = std::min<uint64_t>(N, wave * rows_per_wave);
    const uint64_t r1 = std::min<uint64_t>(N, r0 + rows_per_wave);
    const int col = lane & 31;
    const int kb = 8 * (lane >> 5);
    const bool colok = col < n;
    const int ccol = colok ? col : 0;
    uint32_t *sw = sblk[wv];
    floatx16 acc[NM];
    double facc[NM][16];
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int i = 0; i < 16; i++) {
            acc[m][i] = 0.f;
            facc[m][i] = 0.0;
        }
    int steps = 0;
    // codes of rows s0 + kb + t (t < 8) of this lane's column in the staged block
    auto codes_at = [&](uint32_t s0, uint32_t rows, uint32_t (&c)[8]) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            const uint32_t rl = s0 + (uint32_t)kb + (uint32_t)t;
            c[t] = (colok && rl < rows) ? staged_code<T>(sw, rl * (uint32_t)n + (uint32_t)col) : NONE;
        }
    };
    auto values_at = [&](const uint32_t (&c)[8], Entry (&v)[8]) {
#pragma unroll
        for (int t = 0; t < 8; t++) {
            Entry e{};
            if (c[t] != NONE) e = c[t] < (uint32_t)LSM ? slut[ccol * LSP + c[t]] : glut[(uint64_t)ccol * stride + c[t]];
            v[t] = e;
        }
    };
    // one 16-row step: the MFMAs of every method, and the FP64 flush
    auto mfma_step = [&](const Entry (&v)[8]) {
#pragma unroll
        for (int m = 0; m < NM; m++) {
            bf16x8 hi, lo2;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                uint32_t packed;
                if constexpr (NM == 2) packed = m == 0 ? v[t].x : v[t].y;
                else packed = v[t];
                hi[t] = bf16_from_bits((uint16_t)(packed & 0xFFFF));
                lo2[t] = bf16_from_bits((uint16_t)(packed >> 16));
            }
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, hi, acc[m], 0, 0, 0);
            acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hi, lo2, acc[m], 0, 0, 0);
            if (with_lolo) {
                bf16x8 lq;  // lo2 / 4, exact: lo2 * lq = lo * lo
#pragma unroll
                for (int t = 0; t < 8; t++) lq[t] = (__bf16)((float)lo2[t] * 0.25f);
                acc[m] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lo2, lq, acc[m], 0, 0, 0);
            }
        }
        if (++steps == FLUSH_STEPS) {
            steps = 0;
#pragma unroll
            for (int m = 0; m < NM; m++)
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    facc[m][i] += (double)acc[m][i];
                    acc[m][i] = 0.f;
                }
        }
    };
    uint32_t w[CodeBlock<T>::MAXW];
    if (r0 < r1) fetch_codes<T>(codes, r0 * (uint64_t)n, (uint32_t)(std::min<uint64_t>(RB, r1 - r0) * n), w, lane);
    for (uint64_t rb = r0; rb < r1; rb += RB) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(RB, r1 - rb);
        stage_codes<T>(sw, w, lane);
        __builtin_amdgcn_wave_barrier();
        if (rb + RB < r1)
            fetch_codes<T>(codes, (rb + RB) * (uint64_t)n, (uint32_t)(std::min<uint64_t>(RB, r1 - rb - RB) * n), w,
                           lane);
        if (rows == RB) {
            // software pipeline over the block's steps: the table reads of
            // step q + 1 and the code reads of step q + 2 are in flight
            // while the MFMAs of step q run
            constexpr int NS = RB / 16;
            uint32_t c[8];
            Entry v[8];
            codes_at(0, RB, c);
            values_at(c, v);
            codes_at(16, RB, c);
#pragma unroll
            for (int q = 0; q < NS; q++) {
                Entry vn[8];
                if (q + 1 < NS) values_at(c, vn);
                if (q + 2 < NS) codes_at((uint32_t)(q + 2) * 16u, RB, c);
                mfma_step(v);
                if (q + 1 < NS)
#pragma unroll
                    for (int t = 0; t < 8; t++) v[t] = vn[t];
            }
        } else {
            for (uint32_t s0 = 0; s0 < rows; s0 += 16) {
                uint32_t c[8];
                Entry v[8];
                codes_at(s0, rows, c);
                values_at(c, v);
                mfma_step(v);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int i = 0; i < 16; i++) facc[m][i] += (double)acc[m][i];
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int i = 0; i < 16; i++) partial[((wave * NM + m) * 16 + i) * 64 + lane] = facc[m][i];
}

// FP64 baseline: v_mfma_f64_16x16x4 on the same codes and the same staging,
// FP64 values, NM tables per pass.  partial: [wave][NM][3 blocks (00, 01,
// 11)][4][64].
template <class T, int NM>
__global__ __launch_bounds__(GT) void k_gram_f64(const T *__restrict__ codes, uint64_t N, int n,
                                                  const double *__restrict__ lut, uint32_t stride,
                                                  uint64_t rows_per_wave, double *__restrict__ partial) {
    constexpr int L64 = sizeof(T) == 1 && NM == 1 ? 256 : LSM64;  // 8-bit codes: the whole table
    constexpr int P64 = L64 + 1;
    __shared__ double slut[NM][MAX_COLS * P64];
    __shared__ uint32_t sblk[GT / 64][CodeBlock<T>::WORDS];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    for (int m = 0; m < NM; m++)
        for (int i = threadIdx.x; i < n * L64; i += GT)
            slut[m][(i / L64) * P64 + (i % L64)] =
                lut[(uint64_t)m * n * stride + (uint64_t)(i / L64) * stride + (i % L64)];
    __syncthreads();
    const uint64_t wave = (uint64_t)blockIdx.x * (GT / 64) + wv;
    const uint64_t r0 = std::min<uint64_t>(N, wave * rows_per_wave);
    const uint64_t r1 = std::min<uint64_t>(N, r0 + rows_per_wave);
    const int c0 = lane & 15, c1 = 16 + (lane & 15);
    const int kk = lane >> 4;
    const bool ok0 = c0 < n, ok1 = c1 < n;
    const int k0 = ok0 ? c0 : 0, k1 = ok1 ? c1 : 0;
    uint32_t *sw = sblk[wv];
    doublex4 d00[NM], d01[NM], d11[NM];
#pragma unroll
    for (int m = 0; m < NM; m++) {
        d00[m] = doublex4{0, 0, 0, 0};
        d01[m] = doublex4{0, 0, 0, 0};
        d11[m] = doublex4{0, 0, 0, 0};
    }
    auto value = [&](int m, int cc, uint32_t c) -> double {
        return c < (uint32_t)L64 ? slut[m][cc * P64 + c] : lut[(uint64_t)m * n * stride + (uint64_t)cc * stride + c];
    };
    uint32_t w[CodeBlock<T>::MAXW];
    if (r0 < r1) fetch_codes<T>(codes, r0 * (uint64_t)n, (uint32_t)(std::min<uint64_t>(RB, r1 - r0) * n), w, lane);
    for (uint64_t rb = r0; rb < r1; rb += RB) {
        const uint32_t rows = (uint32_t)std::min<uint64_t>(RB, r1 - rb);
        stage_codes<T>(sw, w, lane);
        __builtin_amdgcn_wave_barrier();
        if (rb + RB < r1)
            fetch_codes<T>(codes, (rb + RB) * (uint64_t)n, (uint32_t)(std::min<uint64_t>(RB, r1 - rb - RB) * n), w,
                           lane);
        for (uint32_t s4 = 0; s4 < rows; s4 += 4) {
            const uint32_t rl = s4 + (uint32_t)kk;
            const bool rv = rl < rows;
            const uint32_t ca = (rv && ok0) ? staged_code<T>(sw, rl * (uint32_t)n + (uint32_t)c0) : 0xFFFFFFFFu;
            const uint32_t cb = (rv && ok1) ? staged_code<T>(sw, rl * (uint32_t)n + (uint32_t)c1) : 0xFFFFFFFFu;
#pragma unroll
            for (int m = 0; m < NM; m++) {
                const double a0 = ca != 0xFFFFFFFFu ? value(m, k0, ca) : 0.0;
                const double a1 = cb != 0xFFFFFFFFu ? value(m, k1, cb) : 0.0;
                d00[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, d00[m], 0, 0, 0);
                d01[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a1, d01[m], 0, 0, 0);
                d11[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, d11[m], 0, 0, 0);
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            partial[(((wave * NM + m) * 3 + 0) * 4 + i) * 64 + lane] = d00[m][i];
            partial[(((wave * NM + m) * 3 + 1) * 4 + i) * 64 + lane] = d01[m][i];
            partial[(((wave * NM + m) * 3 + 2) * 4 + i) * 64 + lane] = d11[m][i];
        }
}

// Sum partials over waves in a fixed order (deterministic), in two levels:
// stage r sums waves [r*chunk, (r+1)*chunk) of every element, then the
// stage sums are added in order.
__global__ void k_sum_partials_stage(const double *__restrict__ partial, uint64_t waves, uint32_t per_wave,
                                     uint64_t chunk, double *__restrict__ stage) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t r = blockIdx.y;
    if (e >= per_wave) return;
    const uint64_t w0 = r * chunk, w1 = std::min<uint64_t>(waves, w0 + chunk);
    double s = 0;
    for (uint64_t w = w0; w < w1; w++) s += partial[w * per_wave + e];
    stage[r * per_wave + e] = s;
}

__global__ void k_sum_partials(const double *__restrict__ partial, uint64_t waves, uint32_t per_wave,
                               double *__restrict__ out) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per_wave) return;
    double s = 0;
    for (uint64_t w = 0; w < waves; w++) s += partial[w * per_wave + e];
    out[e] = s;
}

// Cyclic Jacobi on a symmetric n x n FP64 matrix in one wave (n <= 64):
// lane k owns row k of A and of V; each rotation reads the old rows p, q.
__global__ __launch_bounds__(64) void k_jacobi(double *__restrict__ A, double *__restrict__ V, int n,
                                               int max_sweeps, double *__restrict__ evals) {
    __shared__ double a[64][65];
    __shared__ double v[64][65];
    const int k = threadIdx.x;
    for (int j = 0; j < n; j++)
        if (k < n) {
            a[k][j] = A[k * n + j];
            v[k][j] = (k == j) ? 1.0 : 0.0;
        }
    __syncthreads();
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += a[i][j] * a[i][j];
                if (i != j) off += a[i][j] * a[i][j];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a[p][q];
                if (apq == 0.0) continue;
                const double app = a[p][p], aqq = a[q][q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                double akp = 0, akq = 0, vkp = 0, vkq = 0;
                if (k < n) {
                    akp = a[k][p];
                    akq = a[k][q];
                    vkp = v[k][p];
                    vkq = v[k][q];
                }
                __syncthreads();
                if (k < n) {
                    double nkp = c * akp - s * akq, nkq = s * akp + c * akq;
                    a[k][p] = nkp;
                    a[k][q] = nkq;
                    a[p][k] = nkp;
                    a[q][k] = nkq;
                    v[k][p] = c * vkp - s * vkq;
                    v[k][q] = s * vkp + c * vkq;
                }
                __syncthreads();
                if (k == 0) {
                    a[p][p] = app - t * apq;
                    a[q][q] = aqq + t * apq;
                    a[p][q] = 0.0;
                    a[q][p] = 0.0;
                }
                __syncthreads();
            }
    }
    if (k < n) {
        evals[k] = a[k][k];
        for (int j = 0; j < n; j++) V[k * n + j] = v[k][j];
    }
}

__global__ void k_zero_cells(uint32_t *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Round-robin (parallel) cyclic Jacobi for 64 < n <= JACOBI_RR_MAX on one
// workgroup, A and V in device memory.  Each round rotates m / 2 disjoint
// pairs (m = n rounded up to even; index n is a dummy when n is odd): the
// rotation of pair (p, q) zeroes A[p][q] as the one-wave kernel's does, all
// pairs' column rotations are applied (A J, V J), then all row rotations
// (J^T A).  The pairing follows the circle method: index 0 stays, the
// others shift by one place a round, so a sweep of m - 1 rounds meets
// every pair once.
constexpr int JACOBI_RR_MAX = 1024;

__global__ __launch_bounds__(1024) void k_jacobi_rr(double *__restrict__ A, double *__restrict__ V, int n,
                                                    int max_sweeps, double *__restrict__ evals) {
    __shared__ int ring[JACOBI_RR_MAX];
    __shared__ int pp[JACOBI_RR_MAX / 2], qq[JACOBI_RR_MAX / 2];
    __shared__ double cc[JACOBI_RR_MAX / 2], ss[JACOBI_RR_MAX / 2];
    __shared__ double red[2][16];
    __shared__ int stop;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int m = (n + 1) & ~1, half = m / 2;
    for (int i = tid; i < n * n; i += nt) V[i] = (i / n == i % n) ? 1.0 : 0.0;
    for (int i = tid; i < m; i += nt) ring[i] = i;
    __syncthreads();
    for (int sweep = 0; sweep < max_sweeps; sweep++) {
        // convergence: off-diagonal against total squared mass
        double off = 0, tot = 0;
        for (int i = tid; i < n * n; i += nt) {
            const double x = A[i] * A[i];
            tot += x;
            if (i / n != i % n) off += x;
        }
        for (int d = 32; d > 0; d >>= 1) {
            off += __shfl_xor(off, d, 64);
            tot += __shfl_xor(tot, d, 64);
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = off;
            red[1][tid >> 6] = tot;
        }
        __syncthreads();
        if (tid == 0) {
            double o = 0, t = 0;
            for (int w = 0; w < (nt + 63) / 64; w++) {
                o += red[0][w];
                t += red[1][w];
            }
            stop = (o <= 1e-30 * t || o == 0.0) ? 1 : 0;
        }
        __syncthreads();
        if (stop) break;
        for (int round = 0; round < m - 1; round++) {
            // pairs of this round and their rotations (from the current A)
            for (int i = tid; i < half; i += nt) {
                int p = ring[i], q = ring[m - 1 - i];
                if (p > q) {
                    const int t = p;
                    p = q;
                    q = t;
                }
                double c = 1.0, sn = 0.0;
                if (q < n) {
                    const double apq = A[p * n + q];
                    if (apq != 0.0) {
                        const double app = A[p * n + p], aqq = A[q * n + q];
                        const double theta = (aqq - app) / (2.0 * apq);
                        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(t * t + 1.0);
                        sn = t * c;
                    }
                }
                pp[i] = p;
                qq[i] = q;
                cc[i] = c;
                ss[i] = sn;
            }
            __syncthreads();
            // columns: A <- A J, V <- V J
            for (int e = tid; e < half * n; e += nt) {
                const int i = e / n, k = e - i * n, p = pp[i], q = qq[i];
                if (q >= n || ss[i] == 0.0) continue;
                const double c = cc[i], sn = ss[i];
                const double akp = A[k * n + p], akq = A[k * n + q];
                A[k * n + p] = c * akp - sn * akq;
                A[k * n + q] = sn * akp + c * akq;
                const double vkp = V[k * n + p], vkq = V[k * n + q];
                V[k * n + p] = c * vkp - sn * vkq;
                V[k * n + q] = sn * vkp + c * vkq;
            }
            __syncthreads();
            // rows: A <- J^T A
            for (int e = tid; e < half * n; e += nt) {
                const int i = e / n, k = e - i * n, p = pp[i], q = qq[i];
                if (q >= n || ss[i] == 0.0) continue;
                const double c = cc[i], sn = ss[i];
                const double apk = A[p * n + k], aqk = A[q * n + k];
                A[p * n + k] = c * apk - sn * aqk;
                A[q * n + k] = sn * apk + c * aqk;
            }
            __syncthreads();
            // the pair's off-diagonal entries are zero by construction
            for (int i = tid; i < half; i += nt)
                if (qq[i] < n && ss[i] != 0.0) {
                    A[pp[i] * n + qq[i]] = 0.0;
                    A[qq[i] * n + pp[i]] = 0.0;
                }
            // circle method: slots 1..m-1 shift by one
            int nv = 0;
            if (tid < m && tid > 0) nv = ring[tid == 1 ? m - 1 : tid - 1];
            __syncthreads();
            if (tid < m && tid > 0) ring[tid] = nv;
            __syncthreads();
        }
    }
    for (int k = tid; k < n; k += nt) evals[k] = A[k * n + k];
}

// ---- host -----------------------------------------------------------------
namespace {
uint16_t bf16_bits_rne(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
float bf16_to_float(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}
}  // namespace

int CorrEngine::prepare(const void *d_cells, int cell_bytes, uint64_t N, int n, DevBuf &codes) {
    if (n < 1 || n > MAX_SAMPLES)
        throw Error("invalid argument: sample count must be 1.." + std::to_string(MAX_SAMPLES));
    if (cell_bytes != 4 && cell_bytes != 8) throw Error("internal: count cells must be u32 or u64");
    const uint64_t cells = N * (uint64_t)n;
    auto grid_of = [](uint64_t work) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(4096, (work + 255) / 256)); };
    // largest count: sizes the histogram of the small counts
    DevBuf mx;
    mx.ensure(8);
    k_zero_cells<<<1, 64, 0, stream_>>>(mx.as<uint32_t>(), 2);
    KITE_CHECK_LAUNCH();
    if (cells) {
        if (cell_bytes == 4)
            k_max_cells<uint32_t><<<grid_of(cells), 256, 0, stream_>>>((const uint32_t *)d_cells, cells,
                                                                       mx.as<unsigned long long>());
        else
            k_max_cells<uint64_t><<<grid_of(cells), 256, 0, stream_>>>((const uint64_t *)d_cells, cells,
                                                                       mx.as<unsigned long long>());
        KITE_CHECK_LAUNCH();
    }
    uint64_t maxv = 0;
    KITE_HIP(hipMemcpyAsync(&maxv, mx.ptr, 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
    // counts below `small` have a histogram slot: at most 2^16 per column and
    // 2^26 slots in all; larger counts go through a list
    const uint64_t slot_cap = std::max<uint64_t>(HIST_SMALL, (1ull << 26) / (uint64_t)n);
    const uint32_t small = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({maxv + 1, 65536ull, slot_cap}));
    const uint64_t hslots = (uint64_t)n * small;
    hist_.ensure(hslots * 4 + 16);
    unsigned long long *d_nbig = reinterpret_cast<unsigned long long *>(hist_.as<uint32_t>() + hslots + (hslots & 1));
    k_zero_cells<<<(uint32_t)((hslots + 2 + 255) / 256), 256, 0, stream_>>>(hist_.as<uint32_t>(), hslots + 2 + (hslots & 1));
    KITE_CHECK_LAUNCH();
    if (cells) {
        const dim3 g(grid_of(N * (uint64_t)std::min(n, MAX_COLS)), (unsigned)((n + MAX_COLS - 1) / MAX_COLS));
        if (cell_bytes == 4)
            k_rank_hist<uint32_t><<<g, 256, 0, stream_>>>((const uint32_t *)d_cells, N, (uint32_t)n, small,
                                                         hist_.as<uint32_t>(), d_nbig);
        else
            k_rank_hist<uint64_t><<<g, 256, 0, stream_>>>((const uint64_t *)d_cells, N, (uint32_t)n, small,
                                                         hist_.as<uint32_t>(), d_nbig);
        KITE_CHECK_LAUNCH();
    }
    std::vector<uint32_t> hsmall(hslots);
    uint64_t nbig = 0;
    KITE_HIP(hipMemcpyAsync(hsmall.data(), hist_.ptr, hslots * 4, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipMemcpyAsync(&nbig, d_nbig, 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
    // the large counts, sorted by (column, count)
    std::vector<BigCell> big(nbig);
    if (nbig) {
        DevBuf lst;
        lst.ensure(nbig * sizeof(BigCell) + 8);
        unsigned long long *cur = reinterpret_cast<unsigned long long *>(lst.as<uint8_t>() + nbig * sizeof(BigCell));
        k_zero_cells<<<1, 64, 0, stream_>>>(reinterpret_cast<uint32_t *>(cur), 2);
        KITE_CHECK_LAUNCH();
        const uint32_t g = (uint32_t)((cells + 255) / 256);
        if (cell_bytes == 4)
            k_rank_big<uint32_t><<<g, 256, 0, stream_>>>((const uint32_t *)d_cells, cells, (uint32_t)n, small,
                                                        lst.as<BigCell>(), nbig, cur);
        else
            k_rank_big<uint64_t><<<g, 256, 0, stream_>>>((const uint64_t *)d_cells, cells, (uint32_t)n, small,
                                                        lst.as<BigCell>(), nbig, cur);
        KITE_CHECK_LAUNCH();
        KITE_HIP(hipMemcpyAsync(big.data(), lst.ptr, nbig * sizeof(BigCell), hipMemcpyDeviceToHost, stream_));
        KITE_HIP(hipStreamSynchronize(stream_));
        std::sort(big.begin(), big.end(), [](const BigCell &x, const BigCell &y) {
            return x.col != y.col ? x.col < y.col : x.value < y.value;
        });
    }
    // distinct counts of every column, ascending, with multiplicities
    std::vector<std::vector<uint64_t>> vals(n), mult(n);
    std::vector<uint32_t> map(hslots, 0), big_base(n, 0);
    std::vector<uint64_t> big_off(n + 1, 0), big_vals;
    size_t bi = 0;
    for (int j = 0; j < n; j++) {
        const uint32_t *h = hsmall.data() + (uint64_t)j * small;
        for (uint32_t v = 0; v < small; v++)
            if (h[v]) {
                map[(uint64_t)j * small + v] = (uint32_t)vals[j].size();
                vals[j].push_back(v);
                mult[j].push_back(h[v]);
            }
        big_base[j] = (uint32_t)vals[j].size();
        big_off[j] = big_vals.size();
        for (; bi < big.size() && big[bi].col == (uint64_t)j; bi++) {
            if (vals[j].size() > big_base[j] && vals[j].back() == big[bi].value) {
                mult[j].back()++;
            } else {
                big_vals.push_back(big[bi].value);
                vals[j].push_back(big[bi].value);
                mult[j].push_back(1);
            }
        }
    }
    big_off[n] = big_vals.size();
    uint64_t D = 1;
    for (int j = 0; j < n; j++) D = std::max<uint64_t>(D, vals[j].size());
    if (D > 0xFFFFFFFFull) throw Error("internal: more than 2^32 distinct counts in one column");
    const int cb = D <= 256 ? 1 : (D <= 65536 ? 2 : 4);
    const uint32_t stride = (uint32_t)std::max<uint64_t>(D, HIST_SMALL);
    stride_ = stride;
    code_bytes_ = cb;
    hist_h_.assign((uint64_t)n * stride, 0);
    dense_vals_.assign((uint64_t)n * stride, 0);
    for (int j = 0; j < n; j++)
        for (size_t c = 0; c < vals[j].size(); c++) {
            hist_h_[(uint64_t)j * stride + c] = mult[j][c];
            dense_vals_[(uint64_t)j * stride + c] = vals[j][c];
        }
    // codes on the device
    DevBuf dmap, dboff, dbvals, dbbase;
    dmap.ensure(hslots * 4);
    dboff.ensure((n + 1) * 8);
    dbvals.ensure(std::max<size_t>(big_vals.size(), 1) * 8);
    dbbase.ensure(n * 4);
    KITE_HIP(hipMemcpyAsync(dmap.ptr, map.data(), hslots * 4, hipMemcpyHostToDevice, stream_));
    KITE_HIP(hipMemcpyAsync(dboff.ptr, big_off.data(), (n + 1) * 8, hipMemcpyHostToDevice, stream_));
    if (!big_vals.empty())
        KITE_HIP(hipMemcpyAsync(dbvals.ptr, big_vals.data(), big_vals.size() * 8, hipMemcpyHostToDevice, stream_));
    KITE_HIP(hipMemcpyAsync(dbbase.ptr, big_base.data(), n * 4, hipMemcpyHostToDevice, stream_));
    codes.ensure(std::max<uint64_t>(cells, 1) * cb);
    if (cells) {
        const uint32_t g = (uint32_t)((cells + 255) / 256);
        auto launch = [&](auto cell_tag, auto code_tag) {
            using T = decltype(cell_tag);
            using Cd = decltype(code_tag);
            k_rank_codes<T, Cd><<<g, 256, 0, stream_>>>((const T *)d_cells, cells, (uint32_t)n, small,
                                                        dmap.as<uint32_t>(), dboff.as<uint64_t>(),
                                                        dbvals.as<uint64_t>(), dbbase.as<uint32_t>(), codes.as<Cd>());
        };
        if (cell_bytes == 4) {
            if (cb == 1) launch(uint32_t{}, uint8_t{});
            else if (cb == 2) launch(uint32_t{}, uint16_t{});
            else launch(uint32_t{}, uint32_t{});
        } else {
            if (cb == 1) launch(uint64_t{}, uint8_t{});
            else if (cb == 2) launch(uint64_t{}, uint16_t{});
            else launch(uint64_t{}, uint32_t{});
        }
        KITE_CHECK_LAUNCH();
    }
    KITE_HIP(hipStreamSynchronize(stream_));  // host tables above are locals
    return cb;
}

// Standardized value per (column, code): z = (v - mean) / sd, sd over N.
// Codes are the ranks of the column's distinct counts, so the value of a
// code is ln(1 + count) (Pearson) or the 1-based average-tie rank of the
// count (Spearman): the codes before it cover the smaller counts.
void CorrEngine::make_lut(int method, uint64_t N, int n, std::vector<double> &z, std::vector<double> &sd) {
    const uint32_t stride = stride_;
    z.assign((uint64_t)n * stride, 0.0);
    sd.assign(n, 0.0);
    for (int j = 0; j < n; j++) {
        const uint64_t *h = hist_h_.data() + (uint64_t)j * stride;
        std::vector<double> val(stride, 0.0);
        double cum = 0;
        for (uint32_t c = 0; c < stride; c++) {
            if (method == 0) val[c] = std::log1p((double)dense_vals_[(uint64_t)j * stride + c]);
            else val[c] = cum + ((double)h[c] + 1.0) / 2.0;  // average tie rank, 1-based
            cum += (double)h[c];
        }
        double mean = 0;
        for (uint32_t c = 0; c < stride; c++) mean += (double)h[c] * val[c];
        mean /= (double)N;
        if (method == 1) mean = ((double)N + 1.0) / 2.0;  // exact mean of 1..N ranks
        double var = 0;
        for (uint32_t c = 0; c < stride; c++) {
            double d = val[c] - mean;
            var += (double)h[c] * d * d;
        }
        double s = var > 0 ? std::sqrt(var / (double)N) : 0.0;
        sd[j] = s;
        for (uint32_t c = 0; c < stride; c++) z[(uint64_t)j * stride + c] = s > 0 ? (val[c] - mean) / s : 0.0;
    }
}

void CorrEngine::reduce_partials(uint64_t waves, uint32_t per_wave) {
    const uint64_t R = std::min<uint64_t>(256, std::max<uint64_t>(1, waves));
    const uint64_t chunk = (waves + R - 1) / R;
    stage_.ensure(R * per_wave * 8);
    k_sum_partials_stage<<<dim3((per_wave + 255) / 256, (unsigned)R), 256, 0, stream_>>>(
        partial_.as<double>(), waves, per_wave, chunk, stage_.as<double>());
    KITE_CHECK_LAUNCH();
    k_sum_partials<<<(per_wave + 255) / 256, 256, 0, stream_>>>(stage_.as<double>(), R, per_wave, sum_.as<double>());
    KITE_CHECK_LAUNCH();
}

void CorrEngine::gram(const void *d_codes, int code_bytes, uint64_t N, int n,
                      const std::vector<const std::vector<double> *> &zs, bool fp64, bool with_lolo,
                      std::vector<std::vector<double>> &Gs, float *ms) {
    const uint32_t stride = stride_;
    const int NM = (int)zs.size();
    if (NM < 1 || NM > 2) throw Error("internal: one or two tables per Gram pass");
    // about 12K waves (several rounds of 256 CUs x 12 resident waves, so the
    // last round is a small fraction), each a multiple of 512 rows
    const uint64_t rpw = std::max<uint64_t>(512, ((N + 12287) / 12288 + 511) / 512 * 512);
    const uint64_t waves = std::max<uint64_t>(1, (N + rpw - 1) / rpw);
    const uint64_t wgs = (waves + 3) / 4;
    const uint64_t waves_launched = wgs * 4;
    const size_t tsz = (size_t)n * stride;
    hipEvent_t e0, e1;
    KITE_HIP(hipEventCreate(&e0));
    KITE_HIP(hipEventCreate(&e1));
    std::vector<double> out;
    Gs.assign(NM, std::vector<double>(MAX_COLS * MAX_COLS, 0.0));
    if (!fp64) {
        // entry (column, code) holds the NM methods' packed values side by side
        std::vector<uint32_t> lut(NM * tsz);
        for (int m = 0; m < NM; m++)
            for (size_t i = 0; i < tsz; i++) {
                const double zv = (*zs[m])[i];
                uint16_t hi = bf16_bits_rne((float)zv);
                uint16_t lo = bf16_bits_rne((float)(zv - (double)bf16_to_float(hi)));
                uint16_t lo2 = bf16_bits_rne(2.0f * bf16_to_float(lo));  // exact doubling
                lut[i * NM + m] = (uint32_t)hi | ((uint32_t)lo2 << 16);
            }
        lut_.ensure(lut.size() * 4);
        KITE_HIP(hipMemcpyAsync(lut_.ptr, lut.data(), lut.size() * 4, hipMemcpyHostToDevice, stream_));
        const uint32_t per_wave = NM * 16 * 64;
        partial_.ensure(waves_launched * per_wave * 8);
        sum_.ensure(per_wave * 8);
        KITE_HIP(hipEventRecord(e0, stream_));
        if (NM == 1) {
            if (code_bytes == 1)
                k_gram_bf16<uint8_t, 1><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint8_t *)d_codes, N, n,
                                                                           lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                           partial_.as<double>());
            else if (code_bytes == 2)
                k_gram_bf16<uint16_t, 1><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint16_t *)d_codes, N, n,
                                                                            lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                            partial_.as<double>());
            else
                k_gram_bf16<uint32_t, 1><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint32_t *)d_codes, N, n,
                                                                            lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                            partial_.as<double>());
        } else {
            if (code_bytes == 1)
                k_gram_bf16<uint8_t, 2><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint8_t *)d_codes, N, n,
                                                                           lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                           partial_.as<double>());
            else if (code_bytes == 2)
                k_gram_bf16<uint16_t, 2><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint16_t *)d_codes, N, n,
                                                                            lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                            partial_.as<double>());
            else
                k_gram_bf16<uint32_t, 2><<<(uint32_t)wgs, GT, 0, stream_>>>((const uint32_t *)d_codes, N, n,
                                                                            lut_.as<uint32_t>(), stride, rpw, with_lolo,
                                                                            partial_.as<double>());
        }
        KITE_CHECK_LAUNCH();
        reduce_partials(waves_launched, per_wave);
        KITE_HIP(hipEventRecord(e1, stream_));
        out.resize(per_wave);
        KITE_HIP(hipMemcpyAsync(out.data(), sum_.ptr, per_wave * 8, hipMemcpyDeviceToHost, stream_));
        KITE_HIP(hipStreamSynchronize(stream_));
        // lane l, element i -> (m, n) = (8*(i/4) + 4*(l/32) + i%4, l%32)
        for (int mt = 0; mt < NM; mt++) {
            std::vector<double> A(MAX_COLS * MAX_COLS);
            for (int i = 0; i < 16; i++)
                for (int l = 0; l < 64; l++) {
                    int m = 8 * (i / 4) + 4 * (l / 32) + (i % 4), nn = l % 32;
                    A[m * MAX_COLS + nn] = out[(mt * 16 + i) * 64 + l];
                }
            // the split Gram is the symmetric part of the accumulated products
            for (int i = 0; i < MAX_COLS; i++)
                for (int j = 0; j < MAX_COLS; j++)
                    Gs[mt][i * MAX_COLS + j] = 0.5 * (A[i * MAX_COLS + j] + A[j * MAX_COLS + i]);
        }
    } else {
        std::vector<double> lut(NM * tsz);
        for (int m = 0; m < NM; m++) std::copy(zs[m]->begin(), zs[m]->end(), lut.begin() + m * tsz);
        lut_.ensure(lut.size() * 8);
        KITE_HIP(hipMemcpyAsync(lut_.ptr, lut.data(), lut.size() * 8, hipMemcpyHostToDevice, stream_));
        const uint64_t rpw4 = (rpw + 3) / 4 * 4;
        const uint64_t waves4 = std::max<uint64_t>(1, (N + rpw4 - 1) / rpw4);
        const uint64_t wgs4 = (waves4 + 3) / 4;
        const uint32_t per_wave = NM * 3 * 4 * 64;
        partial_.ensure(wgs4 * 4 * per_wave * 8);
        sum_.ensure(per_wave * 8);
        KITE_HIP(hipEventRecord(e0, stream_));
        if (NM == 1) {
            if (code_bytes == 1)
                k_gram_f64<uint8_t, 1><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint8_t *)d_codes, N, n,
                                                                           lut_.as<double>(), stride, rpw4,
                                                                           partial_.as<double>());
            else if (code_bytes == 2)
                k_gram_f64<uint16_t, 1><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint16_t *)d_codes, N, n,
                                                                            lut_.as<double>(), stride, rpw4,
                                                                            partial_.as<double>());
            else
                k_gram_f64<uint32_t, 1><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint32_t *)d_codes, N, n,
                                                                            lut_.as<double>(), stride, rpw4,
                                                                            partial_.as<double>());
        } else {
            if (code_bytes == 1)
                k_gram_f64<uint8_t, 2><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint8_t *)d_codes, N, n,
                                                                           lut_.as<double>(), stride, rpw4,
                                                                           partial_.as<double>());
            else if (code_bytes == 2)
                k_gram_f64<uint16_t, 2><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint16_t *)d_codes, N, n,
                                                                            lut_.as<double>(), stride, rpw4,
                                                                            partial_.as<double>());
            else
                k_gram_f64<uint32_t, 2><<<(uint32_t)wgs4, GT, 0, stream_>>>((const uint32_t *)d_codes, N, n,
                                                                            lut_.as<double>(), stride, rpw4,
                                                                            partial_.as<double>());
        }
        KITE_CHECK_LAUNCH();
        reduce_partials(wgs4 * 4, per_wave);
        KITE_HIP(hipEventRecord(e1, stream_));
        out.resize(per_wave);
        KITE_HIP(hipMemcpyAsync(out.data(), sum_.ptr, per_wave * 8, hipMemcpyDeviceToHost, stream_));
        KITE_HIP(hipStreamSynchronize(stream_));
        // block b in (00, 01, 11); lane l, element i -> (m, n) = (4*i + l/16, l%16)
        // (the FP64 16x16x4 output layout, measured on gfx950)
        for (int mt = 0; mt < NM; mt++)
            for (int b = 0; b < 3; b++)
                for (int i = 0; i < 4; i++)
                    for (int l = 0; l < 64; l++) {
                        int m = 4 * i + l / 16, nn = l % 16;
                        int bm = b == 2 ? 16 : 0, bn = b == 0 ? 0 : 16;
                        double v = out[((mt * 3 + b) * 4 + i) * 64 + l];
                        Gs[mt][(bm + m) * MAX_COLS + bn + nn] = v;
                        if (b == 1) Gs[mt][(bn + nn) * MAX_COLS + bm + m] = v;
                    }
    }
    KITE_HIP(hipEventElapsedTime(ms, e0, e1));
    KITE_HIP(hipEventDestroy(e0));
    KITE_HIP(hipEventDestroy(e1));
}

void CorrEngine::run(const void *d_codes, uint64_t N, int n, const CorrRequest &req, CorrOutput &out) {
    if (n < 1 || n > MAX_SAMPLES)
        throw Error("invalid argument: sample count must be 1.." + std::to_string(MAX_SAMPLES));
    const int code_bytes = code_bytes_;
    // the requested methods share one pass over the codes
    std::vector<int> methods;
    for (int method = 0; method < 2; method++)
        if (req.methods & (1u << method)) methods.push_back(method);
    if (methods.empty()) return;
    std::vector<std::vector<double>> z(methods.size()), sd(methods.size());
    for (size_t q = 0; q < methods.size(); q++) make_lut(methods[q], N, n, z[q], sd[q]);
    // Gram of every requested table, n x n.  Up to MAX_COLS columns in one
    // 32-column tile; more as pairs of 16-column blocks, each pair's codes
    // gathered into a 32-column matrix for the same kernels.
    const size_t NQ = methods.size();
    std::vector<std::vector<double>> G(NQ, std::vector<double>((size_t)n * n, 0.0)), G64;
    std::vector<float> ms64_each(NQ, 0.f);
    float ms = 0;
    int passes = 0;
    auto tile_gram = [&](const void *dc, const std::vector<int> &cols, bool fp64, int q_only, float *tms) {
        // cols: the tile's columns of the n-column matrix (-1: padding)
        const int tn = (int)cols.size();
        std::vector<std::vector<double>> tz;
        std::vector<size_t> qs;
        for (size_t q = 0; q < NQ; q++)
            if (q_only < 0 || (int)q == q_only) qs.push_back(q);
        for (size_t q : qs) {
            std::vector<double> t((size_t)tn * stride_, 0.0);
            for (int c = 0; c < tn; c++)
                if (cols[c] >= 0)
                    std::copy(z[q].begin() + (size_t)cols[c] * stride_, z[q].begin() + (size_t)(cols[c] + 1) * stride_,
                              t.begin() + (size_t)c * stride_);
            tz.push_back(std::move(t));
        }
        std::vector<const std::vector<double> *> zp;
        for (auto &t : tz) zp.push_back(&t);
        std::vector<std::vector<double>> Gt;
        gram(dc, code_bytes, N, tn, zp, fp64, req.with_lolo && !fp64, Gt, tms);
        return std::make_pair(qs, Gt);
    };
    auto scatter = [&](std::vector<std::vector<double>> &dst, const std::vector<size_t> &qs,
                       const std::vector<std::vector<double>> &Gt, const std::vector<int> &cols,
                       const std::vector<char> &take) {
        const int tn = (int)cols.size();
        for (size_t w = 0; w < qs.size(); w++)
            for (int i = 0; i < tn; i++)
                for (int j = 0; j < tn; j++)
                    if (cols[i] >= 0 && cols[j] >= 0 && take[(size_t)i * tn + j])
                        dst[qs[w]][(size_t)cols[i] * n + cols[j]] = Gt[w][(size_t)i * MAX_COLS + j];
    };
    if (req.fp64_baseline) G64.assign(NQ, std::vector<double>((size_t)n * n, 0.0));
    if (n <= MAX_COLS) {
        std::vector<int> cols(n);
        for (int c = 0; c < n; c++) cols[c] = c;
        std::vector<char> all((size_t)n * n, 1);
        auto r = tile_gram(d_codes, cols, false, -1, &ms);
        scatter(G, r.first, r.second, cols, all);
        passes = 1;
        if (req.fp64_baseline)
            // the FP64 Gram is bound by FP64 MFMA throughput, not by reading
            // the codes: one pass per table is its faster form (measured)
            for (size_t q = 0; q < NQ; q++) {
                auto r64 = tile_gram(d_codes, cols, true, (int)q, &ms64_each[q]);
                scatter(G64, r64.first, r64.second, cols, all);
            }
    } else {
        const int HB = MAX_COLS / 2, nb = (n + HB - 1) / HB;
        DevBuf tmp, dcols;
        tmp.ensure(std::max<uint64_t>(N, 1) * MAX_COLS * code_bytes);
        dcols.ensure(MAX_COLS * sizeof(int));
        std::vector<char> diag_done(nb, 0);
        for (int a = 0; a < nb; a++)
            for (int b = a + 1; b < nb; b++) {
                std::vector<int> cols(MAX_COLS, -1);
                for (int c = 0; c < HB; c++) {
                    if (a * HB + c < n) cols[c] = a * HB + c;
                    if (b * HB + c < n) cols[HB + c] = b * HB + c;
                }
                // this pair gives the (a, b) block; the diagonal blocks come
                // from the first pair that holds them
                std::vector<char> take((size_t)MAX_COLS * MAX_COLS, 0);
                for (int i = 0; i < MAX_COLS; i++)
                    for (int j = 0; j < MAX_COLS; j++) {
                        const bool ia = i < HB, ja = j < HB;
                        if (ia != ja) take[(size_t)i * MAX_COLS + j] = 1;
                        else if (ia && !diag_done[a]) take[(size_t)i * MAX_COLS + j] = 1;
                        else if (!ia && !diag_done[b]) take[(size_t)i * MAX_COLS + j] = 1;
                    }
                KITE_HIP(hipMemcpyAsync(dcols.ptr, cols.data(), MAX_COLS * sizeof(int), hipMemcpyHostToDevice,
                                        stream_));
                const uint64_t cells = N * MAX_COLS;
                if (cells) {
                    const uint32_t g = (uint32_t)((cells + 255) / 256);
                    if (code_bytes == 1)
                        k_gather_cols<uint8_t><<<g, 256, 0, stream_>>>((const uint8_t *)d_codes, N, (uint32_t)n,
                                                                       dcols.as<int>(), tmp.as<uint8_t>());
                    else if (code_bytes == 2)
                        k_gather_cols<uint16_t><<<g, 256, 0, stream_>>>((const uint16_t *)d_codes, N, (uint32_t)n,
                                                                        dcols.as<int>(), tmp.as<uint16_t>());
                    else
                        k_gather_cols<uint32_t><<<g, 256, 0, stream_>>>((const uint32_t *)d_codes, N, (uint32_t)n,
                                                                        dcols.as<int>(), tmp.as<uint32_t>());
                    KITE_CHECK_LAUNCH();
                }
                float t = 0;
                auto r = tile_gram(tmp.ptr, cols, false, -1, &t);
                ms += t;
                scatter(G, r.first, r.second, cols, take);
                passes++;
                if (req.fp64_baseline)
                    for (size_t q = 0; q < NQ; q++) {
                        float t64 = 0;
                        auto r64 = tile_gram(tmp.ptr, cols, true, (int)q, &t64);
                        ms64_each[q] += t64;
                        scatter(G64, r64.first, r64.second, cols, take);
                    }
                diag_done[a] = diag_done[b] = 1;
            }
    }
    float ms64 = 0;
    for (float x : ms64_each) ms64 += x;
    out.gram_ms_total = ms;
    out.fp64_ms_total = ms64;
    out.passes = passes;
    out.fp64_passes = req.fp64_baseline ? (int)NQ * passes : 0;
    auto corr_of = [&](const std::vector<double> &g, std::vector<double> &r) {
        r.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double den = std::sqrt(std::max(0.0, g[(size_t)i * n + i]) * std::max(0.0, g[(size_t)j * n + j]));
                r[(size_t)i * n + j] = den > 0 ? g[(size_t)i * n + j] / den : 0.0;
            }
    };
    for (size_t q = 0; q < NQ; q++) {
        const int method = methods[q];
        corr_of(G[q], method == 0 ? out.pearson : out.spearman);
        if (req.fp64_baseline) corr_of(G64[q], method == 0 ? out.pearson_fp64 : out.spearman_fp64);
        // one pass: its time is shared by the methods it computed
        (method == 0 ? out.gram_ms_pearson : out.gram_ms_spearman) = ms / (float)NQ;
        (method == 0 ? out.fp64_ms_pearson : out.fp64_ms_spearman) = ms64_each[q];
        if (method == 0) {
            // covariance of log1p values (N - 1 divisor): r_ij sigma_i sigma_j N / (N - 1),
            // with the FP64 column deviations and r from the Gram (its diagonal
            // is then the exact FP64 variance)
            out.cov.assign((size_t)n * n, 0.0);
            if (N > 1)
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++)
                        out.cov[(size_t)i * n + j] =
                            out.pearson[(size_t)i * n + j] * sd[q][i] * sd[q][j] * (double)N / (double)(N - 1);
        }
    }
}

void jacobi_device(double *A, double *V, double *evals, int n, hipStream_t s) {
    if (n < 1 || n > JACOBI_RR_MAX)
        throw Error("invalid argument: PCA supports 1.." + std::to_string(JACOBI_RR_MAX) + " samples");
    if (n <= 64) k_jacobi<<<1, 64, 0, s>>>(A, V, n, JACOBI_SWEEPS, evals);
    else k_jacobi_rr<<<1, 1024, 0, s>>>(A, V, n, JACOBI_SWEEPS, evals);
    KITE_CHECK_LAUNCH();
}

void CorrEngine::pca(const std::vector<double> &cov, int n, std::vector<double> &evals, std::vector<double> &evecs) {
    if (n < 1 || n > JACOBI_RR_MAX)
        throw Error("invalid argument: PCA supports 1.." + std::to_string(JACOBI_RR_MAX) + " samples");
    DevBuf a, v, e;
    a.ensure((size_t)n * n * 8);
    v.ensure((size_t)n * n * 8);
    e.ensure((size_t)n * 8);
    KITE_HIP(hipMemcpyAsync(a.ptr, cov.data(), (size_t)n * n * 8, hipMemcpyHostToDevice, stream_));
    jacobi_device(a.as<double>(), v.as<double>(), e.as<double>(), n, stream_);
    evals.resize(n);
    evecs.resize((size_t)n * n);
    KITE_HIP(hipMemcpyAsync(evals.data(), e.ptr, (size_t)n * 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipMemcpyAsync(evecs.data(), v.ptr, (size_t)n * n * 8, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
}

}  // namespace kite
