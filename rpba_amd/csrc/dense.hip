// Batched dense Cholesky of the reduced camera system on the FP64 matrix
// cores (the exact-solve form, solver = 1).  The plan, storage and the
// reason for each kernel's shape are in dense.h.
#include "dense.h"
#include "kernels.h"

#include <algorithm>
#include <chrono>
#include <cstdio>

#include "../../include/rpba.h"

namespace rpba {

namespace {

constexpr int T = kTile;
constexpr int LP = T + 1;  // LDS column stride in doubles (odd: no bank conflicts down a column)
typedef double f64x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_dense_fill(double *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.;
}

// identity on the padding rows n .. nt*T - 1 of every system
__global__ __launch_bounds__(64) void k_dense_pad(DenseArgs a)
{
    const int m = blockIdx.x;
    const int r = a.n[m] + (int)threadIdx.x;
    if (r < a.nt[m] * T) a.A[a.base[m] + (long long)r * a.ld[m] + r] = 1.;
}

// identity systems (the autotuner's input)
__global__ __launch_bounds__(256) void k_dense_eye(DenseArgs a)
{
    const int m = blockIdx.y;
    const int r = blockIdx.x * 256 + (int)threadIdx.x;
    if (r < a.nt[m] * T) a.A[a.base[m] + (long long)r * a.ld[m] + r] = 1.;
}

// diagonal camera blocks (lower part) and the right-hand-side row
__global__ __launch_bounds__(128) void k_dense_diag(DenseArgs a, const int *camsys, const double *Sd, const double *b)
{
    const int k = blockIdx.x, t = threadIdx.x;
    const int m = camsys[k];
    const int ld = a.ld[m];
    const long long kl = 9LL * (k - a.cam0[m]);
    double *A = a.A + a.base[m];
    if (t < 81) {
        const int r = t / 9, c = t - 9 * r;
        if (r >= c) A[(kl + c) * ld + kl + r] = Sd[81LL * k + t];
    } else if (t < 90) {
        const int s = t - 81;
        A[(kl + s) * ld + (long long)a.nt[m] * T] = b[9LL * k + s];
    }
}

// upper block S_kl (k < l) of camera rows k, columns l: lower element
// (row 9 l + j, column 9 k + i) = S_kl[i][j]
__global__ __launch_bounds__(128) void k_dense_off(DenseArgs a, const int *camsys, const double *Su, SuMap sm,
                                                   const int *bk, const int *bl, int nblk)
{
    // workgroups stride over the blocks (a dispatch counts its work-items in 32 bits)
    const int t = threadIdx.x;
    if (t >= 81) return;
    const int i = t / 9, j = t - 9 * i;
    for (int e = blockIdx.x; e < nblk; e += gridDim.x) {
        const int k = bk[e], l = bl[e];
        const int m = camsys[k];
        const int ld = a.ld[m];
        const long long kl = 9LL * (k - a.cam0[m]), ll = 9LL * (l - a.cam0[m]);
        a.A[a.base[m] + (kl + i) * ld + ll + j] = Su[su_at(sm, e, i, j)];
    }
}

// diagonal tile (k, k): unblocked right-looking Cholesky in LDS (column-major,
// odd stride).  Step j: pivot, column scale, then the trailing update with
// thread t on column j+1.. (t & 63) and every fourth row from t >> 6, four
// rows loaded before any is written.  Element order of operations:
// a_rc -= l_rj l_cj for j = 0, 1, ..., then l_rj = a_rj / l_jj, as in the
// textbook column algorithm.
__global__ __launch_bounds__(256) void k_potrf(DenseArgs a, const DenseItem *it, int k)
{
    const DenseItem w = it[blockIdx.x];
    const int ld = a.ld[w.m];
    double *A = a.A + a.base[w.m] + (long long)k * T * ld + (long long)k * T;
    __shared__ double L[T * LP];
    const int t = threadIdx.x, c = t & 63, rg = t >> 6;
    for (int cc = rg; cc < T; cc += 4) L[cc * LP + c] = A[(long long)cc * ld + c];
    __syncthreads();
    for (int j = 0; j < T; ++j) {
        // a pivot <= 0 gives NaN, caught by the step's NaN check
        const double d = sqrt(L[j * LP + j]);
        __syncthreads();
        if (t == j) L[j * LP + j] = d;
        if (t > j && t < T) L[j * LP + t] = L[j * LP + t] / d;
        __syncthreads();
        if (c > j) {
            const double ljc = L[j * LP + c];
            double *Lc = L + c * LP;
            const double *Lj = L + j * LP;
            int r = c + rg;
            for (; r + 12 < T; r += 16) {
                const double a0 = Lj[r], a1 = Lj[r + 4], a2 = Lj[r + 8], a3 = Lj[r + 12];
                const double b0 = Lc[r], b1 = Lc[r + 4], b2 = Lc[r + 8], b3 = Lc[r + 12];
                Lc[r] = b0 - a0 * ljc;
                Lc[r + 4] = b1 - a1 * ljc;
                Lc[r + 8] = b2 - a2 * ljc;
                Lc[r + 12] = b3 - a3 * ljc;
            }
            for (; r < T; r += 4) Lc[r] = Lc[r] - Lj[r] * ljc;
        }
        __syncthreads();
    }
    for (int cc = rg; cc < T; cc += 4)
        if (c >= cc) A[(long long)cc * ld + c] = L[cc * LP + c];
}

// tile (i, k) below the diagonal: X L_kk^T = A_ik by forward substitution,
// lane r holding row r of X in registers, L_kk in LDS (broadcast reads)
__global__ __launch_bounds__(64) void k_trsm(DenseArgs a, const DenseItem *it, int k)
{
    const DenseItem w = it[blockIdx.x];
    const int ld = a.ld[w.m];
    const double *Lkk = a.A + a.base[w.m] + (long long)k * T * ld + (long long)k * T;
    double *X = a.A + a.base[w.m] + (long long)k * T * ld + (long long)w.i * T;
    __shared__ double L[T * LP];
    const int r = threadIdx.x;
    double x[T];
#pragma unroll
    for (int c = 0; c < T; ++c) {
        L[c * LP + r] = Lkk[(long long)c * ld + r];
        x[c] = X[(long long)c * ld + r];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < T; ++j) {
        double s = x[j];
#pragma unroll
        for (int m = 0; m < j; ++m) s -= x[m] * L[m * LP + j];
        x[j] = s / L[j * LP + j];
    }
#pragma unroll
    for (int c = 0; c < T; ++c) X[(long long)c * ld + r] = x[c];
}

// trailing update C_ij -= L_ik L_jk^T (tile row i may be the rhs row) as
// v_mfma_f64_16x16x4f64 steps over K = 64.  Operand layout of that
// instruction: A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in
// lane j + 16 (i % 4), element i / 4.  With A = L_jk (16 rows of block cb) and
// B = L_ik^T (16 rows of block rb), D[i][j] = C[16 rb + j][16 cb + i], so each
// group of 16 lanes updates 16 consecutive rows of one column.  Wave rb takes
// row block rb and the four column blocks.  V = 0 stages both tiles in LDS;
// V = 1 reads the operands straight from memory.  Both issue the same MFMA
// sequence on the same values.
template <int V>
__global__ __launch_bounds__(256) void k_update(DenseArgs a, const DenseItem *it, int nitems, int k)
{
    __shared__ double Ls[V == 0 ? 2 * T * LP : 1];
    const int t = threadIdx.x, rb = t >> 6, lane = t & 63, i16 = lane & 15, q = lane >> 4;
    for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
        const DenseItem d = it[w];
        const int ld = a.ld[d.m];
        double *Ab = a.A + a.base[d.m];
        const double *Li = Ab + (long long)k * T * ld + (long long)d.i * T;
        const double *Lj = Ab + (long long)k * T * ld + (long long)d.j * T;
        double *Cc = Ab + (long long)d.j * T * ld + (long long)d.i * T;
        if (V == 0) {
            __syncthreads();
            for (int e = t; e < T * T; e += 256) {
                const int c = e / T, r = e - c * T;
                Ls[c * LP + r] = Li[(long long)c * ld + r];
                Ls[T * LP + c * LP + r] = Lj[(long long)c * ld + r];
            }
            __syncthreads();
        }
        f64x4 acc[4];
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) acc[cb] = f64x4{0., 0., 0., 0.};
        if (V == 0) {
#pragma unroll 4
            for (int k0 = 0; k0 < T; k0 += 4) {
                const int kk = k0 + q;
                const double bv = Ls[kk * LP + rb * 16 + i16];
                double av[4];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) av[cb] = Ls[T * LP + kk * LP + cb * 16 + i16];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[cb], bv, acc[cb], 0, 0, 0);
            }
        } else {
            // every operand of the item is loaded before the first product, so
            // the wavefront waits for memory once rather than once per K step
            // (80 doubles per lane); the products then run back to back
            double bvv[T / 4], avv[4][T / 4];
#pragma unroll
            for (int s = 0; s < T / 4; ++s) {
                const int kk = 4 * s + q;
                bvv[s] = Li[(long long)kk * ld + rb * 16 + i16];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) avv[cb][s] = Lj[(long long)kk * ld + cb * 16 + i16];
            }
#pragma unroll
            for (int s = 0; s < T / 4; ++s)
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(avv[cb][s], bvv[s], acc[cb], 0, 0, 0);
        }
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double *dst = Cc + (long long)(cb * 16 + q + 4 * v) * ld + rb * 16 + i16;
                *dst = *dst - acc[cb][v];
            }
    }
}

// backward sweep L^T x = y, step over tile kb (from the last tile up): every
// workgroup solves x_kb = L_kb,kb^-T (y_kb - p_kb) in LDS; workgroup j == kb
// writes it out, workgroup j < kb adds L_kb,j^T x_kb to p_j.  p_j collects
// its terms in the fixed order kb = nt-1, nt-2, ...
__global__ __launch_bounds__(256) void k_backward(DenseArgs a, const DenseItem *it)
{
    const DenseItem d = it[blockIdx.x];
    const int m = d.m, kb = d.i, j = d.j, ld = a.ld[m], nt = a.nt[m];
    const double *Ab = a.A + a.base[m];
    double *p = a.p + a.pbase[m];
    __shared__ double L[T * LP];
    __shared__ double z[T];
    const int t = threadIdx.x;
    for (int e = t; e < T * T; e += 256) {
        const int c = e / T, r = e - c * T;
        L[c * LP + r] = Ab[(long long)(kb * T + c) * ld + (long long)kb * T + r];
    }
    if (t < T) z[t] = Ab[(long long)(kb * T + t) * ld + (long long)nt * T] - p[kb * T + t];
    __syncthreads();
    for (int mm = T - 1; mm >= 0; --mm) {
        const double xm = z[mm] / L[mm * LP + mm];
        __syncthreads();
        if (t < mm) z[t] -= L[t * LP + mm] * xm;
        if (t == mm) z[t] = xm;
        __syncthreads();
    }
    if (j == kb) {
        if (t < T) {
            const int r = kb * T + t;
            if (r < a.n[m]) a.sol[9LL * a.cam0[m] + r] = z[t];
        }
        return;
    }
    const int wv = t >> 6, lane = t & 63;
    for (int cc = 0; cc < 16; ++cc) {
        const int c = wv * 16 + cc;
        double v = Ab[(long long)(j * T + c) * ld + (long long)kb * T + lane] * z[lane];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) p[j * T + c] += v;
    }
}

template <class Tp>
int dmalloc(Tp **p, size_t n)
{
    *p = nullptr;
    return hipMalloc((void **)p, sizeof(Tp) * std::max<size_t>(n, 1)) == hipSuccess ? 0 : RPBA_E_NOMEM;
}

template <class Tp>
void dfree(Tp *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

DenseSolver::~DenseSolver() { release(); }

void DenseSolver::release()
{
    dfree(A_); dfree(p_); dfree(base_); dfree(pbase_); dfree(ld_); dfree(nt_); dfree(n_); dfree(cam0_);
    dfree(camsys_); dfree(items_);
    nsys_ = maxnt_ = 0;
    total_ = ptotal_ = 0;
}

int DenseSolver::plan(int nsys, const int *cam0, const int *ncam, int ncam_total)
{
    release();
    nsys_ = nsys;
    ncam_total_ = ncam_total;
    h_nt_.assign(nsys, 0);
    h_n_.assign(nsys, 0);
    h_cam0_.assign(cam0, cam0 + nsys);
    std::vector<long long> hb(nsys), hp(nsys);
    std::vector<int> hld(nsys), hcs(std::max(ncam_total, 1), 0);
    for (int m = 0; m < nsys; ++m) {
        h_n_[m] = 9 * ncam[m];
        h_nt_[m] = (h_n_[m] + T - 1) / T;
        hld[m] = (h_nt_[m] + 1) * T;
        hb[m] = total_;
        hp[m] = ptotal_;
        total_ += (long long)hld[m] * h_nt_[m] * T;
        ptotal_ += (long long)h_nt_[m] * T;
        maxnt_ = std::max(maxnt_, h_nt_[m]);
        for (int k = cam0[m]; k < cam0[m] + ncam[m]; ++k) {
            if (k < 0 || k >= ncam_total) return RPBA_E_INVALID;
            hcs[k] = m;
        }
    }
    if (total_ * 8 > (48LL << 30)) return RPBA_E_UNSUPPORTED;  // dense storage above 48 GB
    // work lists: potrf | trsm | update for every step, then the backward sweep
    std::vector<DenseItem> items;
    potrf_off_.assign(maxnt_ + 1, 0);
    trsm_off_.assign(maxnt_ + 1, 0);
    syrk_off_.assign(maxnt_ + 1, 0);
    back_off_.assign(maxnt_ + 1, 0);
    for (int k = 0; k < maxnt_; ++k) {
        potrf_off_[k] = (long long)items.size();
        for (int m = 0; m < nsys; ++m)
            if (h_nt_[m] > k) items.push_back({m, k, k, 0});
    }
    potrf_off_[maxnt_] = (long long)items.size();
    for (int k = 0; k < maxnt_; ++k) {
        trsm_off_[k] = (long long)items.size();
        for (int m = 0; m < nsys; ++m)
            for (int i = k + 1; i <= h_nt_[m] && h_nt_[m] > k; ++i) items.push_back({m, i, k, 0});
    }
    trsm_off_[maxnt_] = (long long)items.size();
    for (int k = 0; k < maxnt_; ++k) {
        syrk_off_[k] = (long long)items.size();
        // column-major walk: items of one tile column are adjacent, and the
        // persistent grid spreads them round-robin over the XCDs
        for (int m = 0; m < nsys; ++m)
            for (int j = k + 1; j < h_nt_[m]; ++j)
                for (int i = j; i <= h_nt_[m]; ++i) items.push_back({m, i, j, 0});
    }
    syrk_off_[maxnt_] = (long long)items.size();
    for (int s = 0; s < maxnt_; ++s) {
        back_off_[s] = (long long)items.size();
        for (int m = 0; m < nsys; ++m) {
            const int kb = h_nt_[m] - 1 - s;
            for (int j = 0; j <= kb; ++j) items.push_back({m, kb, j, 0});
        }
    }
    back_off_[maxnt_] = (long long)items.size();
    int rc = 0;
    rc |= dmalloc(&A_, (size_t)total_);
    rc |= dmalloc(&p_, (size_t)ptotal_);
    rc |= dmalloc(&base_, nsys);
    rc |= dmalloc(&pbase_, nsys);
    rc |= dmalloc(&ld_, nsys);
    rc |= dmalloc(&nt_, nsys);
    rc |= dmalloc(&n_, nsys);
    rc |= dmalloc(&cam0_, nsys);
    rc |= dmalloc(&camsys_, hcs.size());
    rc |= dmalloc(&items_, items.size());
    if (rc) {
        release();
        return RPBA_E_NOMEM;
    }
    bool ok = true;
    ok &= hipMemcpy(base_, hb.data(), sizeof(long long) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(pbase_, hp.data(), sizeof(long long) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(ld_, hld.data(), sizeof(int) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(nt_, h_nt_.data(), sizeof(int) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(n_, h_n_.data(), sizeof(int) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(cam0_, cam0, sizeof(int) * nsys, hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(camsys_, hcs.data(), sizeof(int) * hcs.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!items.empty())
        ok &= hipMemcpy(items_, items.data(), sizeof(DenseItem) * items.size(), hipMemcpyHostToDevice) == hipSuccess;
    return ok ? 0 : RPBA_E_HIP;
}

DenseArgs DenseSolver::args(double *sol) const
{
    DenseArgs a;
    a.A = A_; a.base = base_; a.ld = ld_; a.nt = nt_; a.n = n_; a.cam0 = cam0_;
    a.p = p_; a.pbase = pbase_; a.sol = sol;
    return a;
}

double DenseSolver::flops() const
{
    double f = 0.;
    for (int m = 0; m < nsys_; ++m) {
        const double n = (double)h_nt_[m] * T;
        f += n * n * n / 3.;
    }
    return f;
}

void DenseSolver::assemble(const double *Sd, const double *Su, const SuMap &sm, const int *blk_k, const int *blk_l,
                           int nblk, const double *b, hipStream_t st)
{
    DenseArgs a = args(nullptr);
    const long long g = std::min<long long>((total_ + 255) / 256, 4096);
    k_dense_fill<<<(int)std::max<long long>(g, 1), 256, 0, st>>>(A_, total_);
    if (nsys_) k_dense_pad<<<nsys_, 64, 0, st>>>(a);
    if (ncam_total_) k_dense_diag<<<ncam_total_, 128, 0, st>>>(a, camsys_, Sd, b);
    if (nblk) k_dense_off<<<std::min(nblk, 1 << 20), 128, 0, st>>>(a, camsys_, Su, sm, blk_k, blk_l, nblk);
}

void DenseSolver::launch_potrf(const DenseArgs &a, int k, hipStream_t st)
{
    const long long n = potrf_off_[k + 1] - potrf_off_[k];
    if (n) k_potrf<<<(int)n, 256, 0, st>>>(a, items_ + potrf_off_[k], k);
}

void DenseSolver::launch_trsm(const DenseArgs &a, int k, hipStream_t st)
{
    const long long n = trsm_off_[k + 1] - trsm_off_[k];
    if (n) k_trsm<<<(int)n, 64, 0, st>>>(a, items_ + trsm_off_[k], k);
}

void DenseSolver::launch_update(const DenseArgs &a, int k, hipStream_t st)
{
    const long long n = syrk_off_[k + 1] - syrk_off_[k];
    if (!n) return;
    const int g = (int)std::min<long long>(n, grid_cap_);
    if (variant_ == 0)
        k_update<0><<<g, 256, 0, st>>>(a, items_ + syrk_off_[k], (int)n, k);
    else
        k_update<1><<<g, 256, 0, st>>>(a, items_ + syrk_off_[k], (int)n, k);
}

void DenseSolver::launch_backward(const DenseArgs &a, hipStream_t st)
{
    const long long g = std::min<long long>((ptotal_ + 255) / 256, 1024);
    k_dense_fill<<<(int)std::max<long long>(g, 1), 256, 0, st>>>(p_, ptotal_);
    for (int s = 0; s < maxnt_; ++s) {
        const long long n = back_off_[s + 1] - back_off_[s];
        if (n) k_backward<<<(int)n, 256, 0, st>>>(a, items_ + back_off_[s]);
    }
}

int DenseSolver::autotune(hipStream_t st)
{
    if (!total_) {
        tuned_ = true;
        return 0;
    }
    double *scratch = nullptr;
    if (dmalloc(&scratch, (size_t)total_)) return RPBA_E_NOMEM;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        dfree(scratch);
        return RPBA_E_HIP;
    }
    double *keep = A_;
    A_ = scratch;
    const int caps[4] = {256, 512, 1024, 2048};
    float best = 1e30f;
    int bv = 0, bc = 512;
    char buf[96];
    tune_log_.clear();
    // first pass warms the code objects; the second is the one that counts
    for (int pass = 0; pass < 2; ++pass)
        for (int v = 0; v < 2; ++v)
            for (int ci = 0; ci < 4; ++ci) {
                variant_ = v;
                grid_cap_ = caps[ci];
                // timings do not depend on the values: factor identity systems
                DenseArgs a = args(nullptr);
                const long long g = std::min<long long>((total_ + 255) / 256, 4096);
                k_dense_fill<<<(int)g, 256, 0, st>>>(scratch, total_);
                k_dense_eye<<<dim3((maxnt_ * T + 255) / 256, nsys_), 256, 0, st>>>(a);
                (void)hipEventRecord(e0, st);
                for (int k = 0; k < maxnt_; ++k) {
                    launch_potrf(a, k, st);
                    launch_trsm(a, k, st);
                    launch_update(a, k, st);
                }
                (void)hipEventRecord(e1, st);
                if (hipEventSynchronize(e1) != hipSuccess) {
                    A_ = keep;
                    dfree(scratch);
                    return RPBA_E_HIP;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (pass == 1) {
                    std::snprintf(buf, sizeof buf, "%s%s/%d:%.3fms", tune_log_.empty() ? "" : " ",
                                  v ? "direct" : "lds", caps[ci], ms);
                    tune_log_ += buf;
                    if (ms < best) {
                        best = ms;
                        bv = v;
                        bc = caps[ci];
                    }
                }
            }
    A_ = keep;
    variant_ = bv;
    grid_cap_ = bc;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dfree(scratch);
    tuned_ = true;
    return 0;
}

std::string DenseSolver::tuning() const
{
    char buf[96];
    std::snprintf(buf, sizeof buf, "chosen %s/%d; ", variant_ ? "direct" : "lds", grid_cap_);
    return std::string(buf) + tune_log_;
}

}  // namespace rpba

// Batched dense solve of independent systems given as full row-major
// matrices (include/rpba.h).  The systems are cut into 9 x 9 camera blocks,
// uploaded as the diagonal and upper blocks the solver assembles from, then
// factored and solved in one batch.
extern "C" int rpba_dense_solve_batch(int32_t nsys, const int32_t *ncam, const double *S, const double *b, double *x,
                                      int32_t variant, int32_t cap)
{
    using namespace rpba;
    if (nsys < 1 || !ncam || !S || !b || !x) return RPBA_E_INVALID;
    std::vector<int> cam0(nsys), nc(nsys);
    int total = 0;
    for (int m = 0; m < nsys; ++m) {
        if (ncam[m] < 1) return RPBA_E_INVALID;
        cam0[m] = total;
        nc[m] = ncam[m];
        total += ncam[m];
    }
    std::vector<double> Sd(81 * (size_t)total), Su;
    std::vector<int> bk, bl;
    size_t soff = 0;
    // one block per slab (slot b, slab b): element (i, j) of block b at 81 b + 9 j + i
    long long nbt = 0;
    for (int m = 0; m < nsys; ++m) nbt += (long long)nc[m] * (nc[m] - 1) / 2;
    Su.assign(81 * (size_t)std::max(nbt, 1LL), 0.);
    for (int m = 0; m < nsys; ++m) {
        const size_t n = 9 * (size_t)nc[m];
        const double *Sm = S + soff;
        for (int k = 0; k < nc[m]; ++k)
            for (int l = k; l < nc[m]; ++l) {
                if (k == l) {
                    double *dst = Sd.data() + 81 * (size_t)(cam0[m] + k);
                    for (int i = 0; i < 9; ++i)
                        for (int j = 0; j < 9; ++j) dst[9 * i + j] = Sm[(9 * (size_t)k + i) * n + 9 * (size_t)l + j];
                } else {
                    const size_t e = bk.size();
                    bk.push_back(cam0[m] + k);
                    bl.push_back(cam0[m] + l);
                    for (int i = 0; i < 9; ++i)
                        for (int j = 0; j < 9; ++j)
                            Su[81 * e + 9 * j + i] = Sm[(9 * (size_t)k + i) * n + 9 * (size_t)l + j];
                }
            }
        soff += n * n;
    }
    const size_t nb = bk.size(), n9 = 9 * (size_t)total;
    DenseSolver ds;
    int rc = ds.plan(nsys, cam0.data(), nc.data(), total);
    if (rc) return rc;
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return RPBA_E_HIP;
    double *dSd = nullptr, *dSu = nullptr, *db = nullptr, *dx = nullptr;
    int *dk = nullptr, *dl = nullptr;
    rc |= dmalloc(&dSd, Sd.size());
    rc |= dmalloc(&dSu, Su.size());
    rc |= dmalloc(&db, n9);
    rc |= dmalloc(&dx, n9);
    rc |= dmalloc(&dk, nb);
    rc |= dmalloc(&dl, nb);
    int *dio = nullptr;  // 0, 1, ..., nb: slots, slabs and slab starts of one block per slab
    rc |= dmalloc(&dio, nb + 1);
    bool ok = rc == 0;
    if (ok) {
        ok &= hipMemcpy(dSd, Sd.data(), sizeof(double) * Sd.size(), hipMemcpyHostToDevice) == hipSuccess;
        if (nb) {
            ok &= hipMemcpy(dSu, Su.data(), sizeof(double) * Su.size(), hipMemcpyHostToDevice) == hipSuccess;
            ok &= hipMemcpy(dk, bk.data(), sizeof(int) * nb, hipMemcpyHostToDevice) == hipSuccess;
            ok &= hipMemcpy(dl, bl.data(), sizeof(int) * nb, hipMemcpyHostToDevice) == hipSuccess;
            std::vector<int> io(nb + 1);
            for (size_t e = 0; e <= nb; ++e) io[e] = (int)e;
            ok &= hipMemcpy(dio, io.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice) == hipSuccess;
        }
        ok &= hipMemcpy(db, b, sizeof(double) * n9, hipMemcpyHostToDevice) == hipSuccess;
    }
    if (ok) {
        if (variant < 0)
            ok = ds.autotune(st) == 0;
        else
            ds.set_variant(variant, cap);
    }
    if (ok) {
        const SuMap sm{dio, dio, dio};
        ds.assemble(dSd, dSu, sm, dk, dl, (int)nb, db, st);
        ds.factor_solve(dx, st, [](const char *, auto f) { f(); });
        ok = hipStreamSynchronize(st) == hipSuccess && hipGetLastError() == hipSuccess;
        if (ok) ok = hipMemcpy(x, dx, sizeof(double) * n9, hipMemcpyDeviceToHost) == hipSuccess;
    }
    dfree(dSd); dfree(dSu); dfree(db); dfree(dx); dfree(dk); dfree(dl); dfree(dio);
    (void)hipStreamDestroy(st);
    if (!ok) return rc ? rc : RPBA_E_HIP;
    return RPBA_OK;
}
