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
    float gram_ms_pearson = 0, gram_ms_spearman = 0;   // split-bf16 Gram (+ reduce), share of the pass
    float fp64_ms_pearson = 0, fp64_ms_spearman = 0;   // FP64 Gram (+ reduce), share of the pass
    float gram_ms_total = 0, fp64_ms_total = 0;        // the whole pass (all requested methods)
    int passes = 0;                                    // passes over the codes of the split-bf16 Gram
    int fp64_passes = 0;                               // passes of the FP64 Gram (one per table)
};

// Cyclic Jacobi sweeps of the PCA eigendecomposition.
constexpr int JACOBI_SWEEPS = 60;

// Eigendecomposition of the symmetric n x n FP64 matrix A (device, row
// major, n <= 64, overwritten) on one wave: eigenvectors into the columns of
// V (n x n), eigenvalues (unsorted) into evals.  Queued on stream s.
void jacobi_device(double *A, double *V, double *evals, int n, hipStream_t s);

// Columns (samples) the correlation takes.
constexpr int MAX_SAMPLES = 4096;

class CorrEngine {
   public:
    explicit CorrEngine(hipStream_t s) : stream_(s) {}
    // Rank codes of a count matrix (N x n row major, u32 or u64 cells,
    // device): every cell becomes the rank of its count among the distinct
    // counts of its column, 1, 2 or 4 bytes wide (returned); the value
    // tables behind the codes are kept for run().
    int prepare(const void *d_cells, int cell_bytes, uint64_t N, int n, DevBuf &codes);
    // Correlation of the codes prepare() made.
    void run(const void *d_codes, uint64_t N, int n, const CorrRequest &req, CorrOutput &out);
    void pca(const std::vector<double> &cov, int n, std::vector<double> &evals, std::vector<double> &evecs);

   private:
    void make_lut(int method, uint64_t N, int n, std::vector<double> &z, std::vector<double> &sd);
    // One pass over the codes (n <= 32 columns) computes the Gram of every
    // table in zs (1 or 2); Gs come back in a 32 x 32 layout.
    void gram(const void *d_codes, int code_bytes, uint64_t N, int n, const std::vector<const std::vector<double> *> &zs,
              bool fp64, bool with_lolo, std::vector<std::vector<double>> &Gs, float *ms);

    void reduce_partials(uint64_t waves, uint32_t per_wave);

    hipStream_t stream_;
    uint32_t stride_ = 0;   // value-table entries per column (>= 256)
    int code_bytes_ = 1;
    DevBuf hist_, lut_, partial_, sum_, stage_;
    std::vector<uint64_t> hist_h_;      // [column][code] -> multiplicity
    std::vector<uint64_t> dense_vals_;  // [column][code] -> count value
};

}  // namespace kite
