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

template <class T>
__global__ __launch_bounds__(256) void k_col_hist(const T *__restrict__ codes, uint64_t N, int n,
                                                  uint32_t stride, uint32_t *__restrict__ hist) {
    __shared__ uint32_t sm[MAX_COLS * HIST_SMALL];
    for (int i = threadIdx.x; i < n * HIST_SMALL; i += 256) sm[i] = 0;
    __syncthreads();
    const uint64_t total = N * (uint64_t)n;
    for (uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (uint64_t)gridDim.x * 256) {
        uint32_t c = (uint32_t)codes[e];
        uint32_t j = (uint32_t)(e % (uint64_t)n);
        if (c < HIST_SMALL) atomicAdd(&sm[j * HIST_SMALL + c], 1u);
        else atomicAdd(&hist[(uint64_t)j * stride + c], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n * HIST_SMALL; i += 256) {
        uint32_t v = sm[i];
        if (v) atomicAdd(&hist[(uint64_t)(i / HIST_SMALL) * stride + (i % HIST_SMALL)], v);
    }
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

__global__ void k_max_u32(const uint32_t *__restrict__ p, uint64_t n, uint32_t *__restrict__ out) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        m = std::max(m, p[i]);
    for (int d = 32; d > 0; d >>= 1) m = std::max(m, (uint32_t)__shfl_xor((int)m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_zero_cells(uint32_t *p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// 8-bit codes: each count becomes the rank of its value among the distinct
// values of its column (map: [column][stride] u8).
__global__ void k_dense_codes(const uint32_t *__restrict__ in, uint64_t cells, uint32_t n,
                              const uint8_t *__restrict__ map, uint32_t stride, uint8_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cells) out[i] = map[(uint64_t)(i % n) * stride + in[i]];
}

template <class T>
__global__ void k_u32_to_codes(const uint32_t *__restrict__ in, uint64_t n, T *__restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[i];
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

void CorrEngine::column_stats(const void *d_codes, int code_bytes, uint64_t N, int n, uint32_t maxc) {
    const uint32_t stride = std::max<uint32_t>(maxc + 1, HIST_SMALL);
    stride_ = stride;
    hist_.ensure((uint64_t)n * stride * 4);
    {
        const uint64_t cells = (uint64_t)n * stride;
        k_zero_cells<<<(uint32_t)((cells + 255) / 256), 256, 0, stream_>>>(hist_.as<uint32_t>(), cells);
        KITE_CHECK_LAUNCH();
    }
    const uint64_t total = N * (uint64_t)n;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(4096, std::max<uint64_t>(1, (total + 255) / 256));
    if (total) {
        if (code_bytes == 2)
            k_col_hist<uint16_t><<<grid, 256, 0, stream_>>>((const uint16_t *)d_codes, N, n, stride, hist_.as<uint32_t>());
        else
            k_col_hist<uint32_t><<<grid, 256, 0, stream_>>>((const uint32_t *)d_codes, N, n, stride, hist_.as<uint32_t>());
        KITE_CHECK_LAUNCH();
    }
    hist_h_.resize((uint64_t)n * stride);
    KITE_HIP(hipMemcpyAsync(hist_h_.data(), hist_.ptr, hist_h_.size() * 4, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
}

// Standardized value per (column, count): z = (v - mean) / sd, sd over N.
void CorrEngine::make_lut(int method, uint64_t N, int n, std::vector<double> &z, std::vector<double> &sd) {
    const uint32_t stride = stride_;
    z.assign((uint64_t)n * stride, 0.0);
    sd.assign(n, 0.0);
    for (int j = 0; j < n; j++) {
        const uint32_t *h = hist_h_.data() + (uint64_t)j * stride;
        std::vector<double> val(stride, 0.0);
        double cum = 0;
        for (uint32_t c = 0; c < stride; c++) {
            if (method == 0) val[c] = std::log1p((double)(dense_ ? dense_vals_[(uint64_t)j * stride + c] : c));
            else val[c] = cum + ((double)h[c] + 1.0) / 2.0;  // average tie rank, 1-based
            cum += h[c];
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

void CorrEngine::run(const void *d_codes, int code_bytes, uint64_t N, int n, uint32_t maxc, const CorrRequest &req,
                     CorrOutput &out) {
    if (n < 1 || n > MAX_COLS) throw Error("invalid argument: sample count must be 1.." + std::to_string(MAX_COLS));
    if (code_bytes == 1) {
        if (!dense_) throw Error("internal: 8-bit codes without their value table");
    } else {
        dense_ = false;
        column_stats(d_codes, code_bytes, N, n, maxc);
    }
    // the requested methods share one pass over the codes
    std::vector<int> methods;
    for (int method = 0; method < 2; method++)
        if (req.methods & (1u << method)) methods.push_back(method);
    if (methods.empty()) return;
    std::vector<std::vector<double>> z(methods.size()), sd(methods.size());
    std::vector<const std::vector<double> *> zp;
    for (size_t q = 0; q < methods.size(); q++) {
        make_lut(methods[q], N, n, z[q], sd[q]);
        zp.push_back(&z[q]);
    }
    std::vector<std::vector<double>> G, G64;
    float ms = 0, ms64 = 0;
    gram(d_codes, code_bytes, N, n, zp, false, req.with_lolo, G, &ms);
    std::vector<float> ms64_each(methods.size(), 0.f);
    if (req.fp64_baseline) {
        // the FP64 Gram is bound by FP64 MFMA throughput, not by reading the
        // codes: one pass per table is its faster form (measured), so the
        // baseline takes that form
        G64.resize(methods.size());
        for (size_t q = 0; q < methods.size(); q++) {
            std::vector<std::vector<double>> g1;
            gram(d_codes, code_bytes, N, n, {zp[q]}, true, false, g1, &ms64_each[q]);
            G64[q] = std::move(g1[0]);
            ms64 += ms64_each[q];
        }
    }
    out.gram_ms_total = ms;
    out.fp64_ms_total = ms64;
    out.passes = 1;
    out.fp64_passes = req.fp64_baseline ? (int)methods.size() : 0;
    auto corr_of = [&](const std::vector<double> &g, std::vector<double> &r) {
        r.assign((size_t)n * n, 0.0);
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double den = std::sqrt(std::max(0.0, g[i * MAX_COLS + i]) * std::max(0.0, g[j * MAX_COLS + j]));
                r[i * n + j] = den > 0 ? g[i * MAX_COLS + j] / den : 0.0;
            }
    };
    for (size_t q = 0; q < methods.size(); q++) {
        const int method = methods[q];
        corr_of(G[q], method == 0 ? out.pearson : out.spearman);
        if (req.fp64_baseline) corr_of(G64[q], method == 0 ? out.pearson_fp64 : out.spearman_fp64);
        // one pass: its time is shared by the methods it computed
        (method == 0 ? out.gram_ms_pearson : out.gram_ms_spearman) = ms / (float)methods.size();
        (method == 0 ? out.fp64_ms_pearson : out.fp64_ms_spearman) = ms64_each[q];
        if (method == 0) {
            // covariance of log1p values (N - 1 divisor) from the standardized Gram
            out.cov.assign((size_t)n * n, 0.0);
            if (N > 1)
                for (int i = 0; i < n; i++)
                    for (int j = 0; j < n; j++)
                        out.cov[i * n + j] = G[q][i * MAX_COLS + j] * sd[q][i] * sd[q][j] / (double)(N - 1);
        }
    }
}

void jacobi_device(double *A, double *V, double *evals, int n, hipStream_t s) {
    if (n < 1 || n > 64) throw Error("invalid argument: PCA supports 1..64 samples");
    k_jacobi<<<1, 64, 0, s>>>(A, V, n, JACOBI_SWEEPS, evals);
    KITE_CHECK_LAUNCH();
}

void CorrEngine::pca(const std::vector<double> &cov, int n, std::vector<double> &evals, std::vector<double> &evecs) {
    if (n < 1 || n > 64) throw Error("invalid argument: PCA supports 1..64 samples");
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

uint32_t CorrEngine::max_count(const uint32_t *d_counts, uint64_t cells) {
    DevBuf m;
    m.ensure(4);
    k_zero_cells<<<1, 64, 0, stream_>>>(m.as<uint32_t>(), 1);
    KITE_CHECK_LAUNCH();
    if (cells) {
        uint32_t g = (uint32_t)std::min<uint64_t>(2048, (cells + 255) / 256);
        k_max_u32<<<g, 256, 0, stream_>>>(d_counts, cells, m.as<uint32_t>());
        KITE_CHECK_LAUNCH();
    }
    uint32_t h = 0;
    KITE_HIP(hipMemcpyAsync(&h, m.ptr, 4, hipMemcpyDeviceToHost, stream_));
    KITE_HIP(hipStreamSynchronize(stream_));
    return h;
}

bool CorrEngine::dense_codes(const uint32_t *d_counts, uint64_t N, int n, uint32_t maxc, DevBuf &codes) {
    dense_ = false;
    if (maxc >= (1u << 20) || n < 1 || n > MAX_COLS) return false;
    const uint64_t cells = N * (uint64_t)n;
    // histogram of the raw counts (the value table needs it anyway)
    column_stats(d_counts, 4, N, n, maxc);
    const uint32_t stride0 = stride_;
    std::vector<uint8_t> map((uint64_t)n * stride0, 0);
    std::vector<uint32_t> vals((uint64_t)n * 256, 0), hcode((uint64_t)n * 256, 0);
    for (int j = 0; j < n; j++) {
        const uint32_t *h = hist_h_.data() + (uint64_t)j * stride0;
        uint32_t code = 0;
        for (uint32_t v = 0; v < stride0; v++) {
            if (!h[v]) continue;
            if (code == 256) return false;  // more than 256 distinct counts in this column
            map[(uint64_t)j * stride0 + v] = (uint8_t)code;
            vals[(uint64_t)j * 256 + code] = v;
            hcode[(uint64_t)j * 256 + code] = h[v];
            code++;
        }
    }
    DevBuf dmap;
    dmap.ensure(map.size());
    KITE_HIP(hipMemcpyAsync(dmap.ptr, map.data(), map.size(), hipMemcpyHostToDevice, stream_));
    codes.ensure(std::max<uint64_t>(cells, 1));
    if (cells) {
        k_dense_codes<<<(uint32_t)((cells + 255) / 256), 256, 0, stream_>>>(d_counts, cells, (uint32_t)n,
                                                                            dmap.as<uint8_t>(), stride0,
                                                                            codes.as<uint8_t>());
        KITE_CHECK_LAUNCH();
    }
    KITE_HIP(hipStreamSynchronize(stream_));  // `map` is a local vector
    // from here on the value table is indexed by code
    hist_h_.swap(hcode);
    dense_vals_.swap(vals);
    stride_ = 256;
    dense_ = true;
    return true;
}

void CorrEngine::to_codes(const uint32_t *d_counts, uint64_t cells, DevBuf &codes, int code_bytes) {
    codes.ensure(std::max<uint64_t>(cells, 1) * code_bytes);
    if (!cells) return;
    const uint32_t g = (uint32_t)((cells + 255) / 256);
    if (code_bytes == 2) k_u32_to_codes<uint16_t><<<g, 256, 0, stream_>>>(d_counts, cells, codes.as<uint16_t>());
    else k_u32_to_codes<uint32_t><<<g, 256, 0, stream_>>>(d_counts, cells, codes.as<uint32_t>());
    KITE_CHECK_LAUNCH();
}

}  // namespace kite
