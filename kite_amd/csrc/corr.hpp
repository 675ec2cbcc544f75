// corr.hpp -- correlation (Pearson on log1p counts, Spearman on average
// ranks) of union count columns via an MFMA Gram, and the FP64 Jacobi PCA.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace kite {

struct CorrRequest {
    uint32_t methods = 3;       // bit 0 Pearson, bit 1 Spearman
    bool fp64_baseline = false; // also run the FP64 MFMA Gram
    bool with_lolo = false;     // add the lo*lo product
};

struct CorrOutput {
    std::vector<double> pearson, spearman, cov;        // n x n
    std::vector<double> pearson_fp64, spearman_fp64;   // FP64 Gram results
    float gram_ms_pearson = 0, gram_ms_spearman = 0;   // split-bf16 Gram (+ reduce)
    float fp64_ms_pearson = 0, fp64_ms_spearman = 0;   // FP64 Gram (+ reduce)
};

class CorrEngine {
   public:
    explicit CorrEngine(hipStream_t s) : stream_(s) {}
    // d_codes: N x n row-major count values (u16 or u32), all <= maxc.
    void run(const void *d_codes, int code_bytes, uint64_t N, int n, uint32_t maxc, const CorrRequest &req,
             CorrOutput &out);
    void pca(const std::vector<double> &cov, int n, std::vector<double> &evals, std::vector<double> &evecs);
    void to_codes(const uint32_t *d_counts, uint64_t cells, DevBuf &codes, int code_bytes);
    uint32_t max_count(const uint32_t *d_counts, uint64_t cells);

   private:
    void column_stats(const void *d_codes, int code_bytes, uint64_t N, int n, uint32_t maxc);
    void make_lut(int method, uint64_t N, int n, std::vector<double> &z, std::vector<double> &sd);
    void gram(const void *d_codes, int code_bytes, uint64_t N, int n, const std::vector<double> &z, bool fp64,
              bool with_lolo, std::vector<double> &G, float *ms);

    void reduce_partials(uint64_t waves, uint32_t per_wave);

    hipStream_t stream_;
    uint32_t stride_ = 0;
    DevBuf hist_, lut_, partial_, sum_, stage_;
    std::vector<uint32_t> hist_h_;
};

}  // namespace kite
