// Dense Cholesky solve of the reduced camera system on the FP64 matrix cores,
// batched over many independent systems of uneven size.
//
// The reference solves S x = b with block-Jacobi PCG (SparsePCG,
// rbacore.cpp:1807-1914).  This is the exact-solve form: S (9 slots per
// camera; fixed slots are identity rows with zero right-hand side) is
// assembled densely, factored S = L L^T by a right-looking tiled Cholesky
// whose trailing updates run as v_mfma_f64_16x16x4f64 products, and solved by
// two triangular sweeps.  The right-hand side rides along as one extra row of
// tiles below the matrix, so the factorisation leaves y = L^-1 b in that row
// and only the backward sweep L^T x = y remains.
//
// One launch per kernel and step covers every system at once: the work lists
// hold (system, tile row, tile column) items for all systems, so systems of
// different sizes share launches and the small ones simply run out of items
// early.  The trailing-update launch is persistent (a capped grid walks the
// item list), and its grid cap and operand path (LDS-staged or direct) are
// chosen by a timing autotuner.  Every variant computes each element with the
// same instruction sequence, so the tuned choice never changes the result.
#pragma once

#include <hip/hip_runtime.h>
#include "kernels.h"

#include <cstdint>
#include <string>
#include <vector>

namespace rpba {

constexpr int kTile = 64;  // Cholesky tile edge

struct DenseItem {
    int m, i, j, pad;  // system, tile row, tile column
};

struct DenseArgs {
    double *A;                 // all systems; system m at A + base[m], column-major, leading dimension ld[m]
    const long long *base;
    const int *ld, *nt, *n;    // leading dimension, tiles of the system (the rhs is tile row nt), true dimension
    const int *cam0;           // first camera of system m (rows 9 (k - cam0) + s)
    double *p;                 // backward-sweep partial sums, system m at p + pbase[m] (nt tiles)
    const long long *pbase;
    double *sol;               // solution, 9 per camera (global camera index)
};

// Plan and storage of a batch of dense systems.  Each system m covers the
// cameras [cam0[m], cam0[m] + ncam[m]); a camera belongs to one system.
class DenseSolver {
public:
    ~DenseSolver();
    // sizes and work lists; allocates device storage.  Returns 0 or a
    // negative rpba_status.
    int plan(int nsys, const int *cam0, const int *ncam, int ncam_total);
    // zero the systems, then scatter S_kk (81 per camera, row-major), the
    // upper blocks S_kl (k < l, row-major 9 x 9, block list blk_k/blk_l) and
    // b (9 per camera) into the lower triangles and right-hand-side rows.
    void assemble(const double *Sd, const double *Su, const SuMap &sm, const int *blk_k, const int *blk_l, int nblk,
                  const double *b, hipStream_t st);
    // factor and solve; writes sol (9 per camera of the planned systems).
    // timer(name, launch) wraps each launch (for per-kernel HIP-event timing).
    template <class F>
    void factor_solve(double *sol, hipStream_t st, F timer)
    {
        DenseArgs a = args(sol);
        for (int k = 0; k < maxnt_; ++k) {
            timer("chol_potrf", [&] { launch_potrf(a, k, st); });
            timer("chol_trsm", [&] { launch_trsm(a, k, st); });
            if (syrk_off_[k + 1] > syrk_off_[k]) timer("chol_update", [&] { launch_update(a, k, st); });
        }
        timer("chol_solve", [&] { launch_backward(a, st); });
    }
    // time every trailing-update variant on identity systems of the planned
    // sizes (in a scratch buffer) and keep the fastest; the choice does not
    // change results.  Returns 0 or a negative rpba_status.
    int autotune(hipStream_t st);
    bool tuned() const { return tuned_; }
    void set_variant(int v, int cap)
    {
        variant_ = v ? 1 : 0;
        grid_cap_ = cap > 0 ? cap : 512;
        tuned_ = true;
    }
    std::string tuning() const;  // chosen variant and the timings, one line
    long long matrix_doubles() const { return total_; }
    int max_tiles() const { return maxnt_; }
    double flops() const;        // factorisation flops (sum of n^3 / 3 over the padded systems)

private:
    DenseArgs args(double *sol) const;
    void launch_potrf(const DenseArgs &a, int k, hipStream_t st);
    void launch_trsm(const DenseArgs &a, int k, hipStream_t st);
    void launch_update(const DenseArgs &a, int k, hipStream_t st);
    void launch_backward(const DenseArgs &a, hipStream_t st);
    void release();

    int nsys_ = 0, maxnt_ = 0, ncam_total_ = 0;
    long long total_ = 0, ptotal_ = 0;
    std::vector<int> h_nt_, h_n_, h_cam0_;
    std::vector<long long> potrf_off_, trsm_off_, syrk_off_, back_off_;
    double *A_ = nullptr, *p_ = nullptr;
    long long *base_ = nullptr, *pbase_ = nullptr;
    int *ld_ = nullptr, *nt_ = nullptr, *n_ = nullptr, *cam0_ = nullptr, *camsys_ = nullptr;
    DenseItem *items_ = nullptr;
    // trailing-update variant: operand path (0 = LDS-staged, 1 = direct) and grid cap
    int variant_ = 0, grid_cap_ = 512;
    bool tuned_ = false;
    std::string tune_log_;
};

}  // namespace rpba
