// Kernels of the partitioned solve: CalcLSAXMulti per point (thread per
// point) and the block-segmented PCG over the union of the parts' reduced
// camera systems.  Every reduction runs in a fixed order; no floating-point
// atomics.
#include "partition.h"

namespace rpba {

namespace {

__device__ __forceinline__ double pwave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// block-wide sum in a fixed order; every thread gets it.  sh: blockDim/64
__device__ __forceinline__ double pblock_sum(double v, double *sh)
{
    v = pwave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.;
    for (int i = 0; i < nw; ++i) s += sh[i];
    return s;
}

// 3x3 inverse by cofactors (Eigen's Matrix3d::inverse, rbacore.cpp:1344)
__device__ __forceinline__ void pinv3(const double m[9], double r[9])
{
    double cf[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cf[i * 3 + j] = m[i1 * 3 + j1] * m[i2 * 3 + j2] - m[i1 * 3 + j2] * m[i2 * 3 + j1];
        }
    const double det = cf[0] * m[0] + cf[3] * m[3] + cf[6] * m[6];
    const double id = 1. / det;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) r[i * 3 + j] = cf[j * 3 + i] * id;
}

// one entry's projection at the float point (CalcCorrPointMulti,
// rpbacore.cpp:797-830, RadialReconstructionImage rpba.h:850-857 with the
// float Camera values): the float normalised coordinates and 1/z the
// derivatives use, and the residual
struct LProj {
    float p1v, p2v, p3v;
    double dx, dy;
};

__device__ __forceinline__ LProj lsax_proj(const double *PM, float f, float d0, float d1, float ox, float oy, double X0,
                                           double X1, double X2)
{
    LProj r;
    double p1 = PM[0] * X0 + PM[1] * X1 + PM[2] * X2 + PM[3];
    double p2 = PM[4] * X0 + PM[5] * X1 + PM[6] * X2 + PM[7];
    double p3 = PM[8] * X0 + PM[9] * X1 + PM[10] * X2 + PM[11];
    if (fabs(p3) < 1.e-3)
        p3 = (p3 > 0.) ? 1.e3 : -1.e3;
    else
        p3 = 1. / p3;
    p1 *= p3;
    p2 *= p3;
    r.p1v = (float)p1;
    r.p2v = (float)p2;
    r.p3v = (float)p3;
    const double dist = p1 * p1 + p2 * p2;
    const double z = 1. + (double)((double)d0 * dist + (double)d1 * dist * dist);
    p1 *= z;
    p2 *= z;
    p1 *= (double)f;
    p2 *= (double)f;
    r.dx = (double)ox - p1;
    r.dy = (double)oy - p2;
    return r;
}

// the calibrated point derivatives of one entry from its float projection
// (rpbacore.cpp:611-629 with CalcCalibratedDerivatives, 841-877; skew 0)
__device__ __forceinline__ void lsax_jac(const double *PM, float f, float d0, float d1, const LProj &p, double B[2][3])
{
    const double x = p.p1v, y = p.p2v;
    const double P3 = p.p3v, P1 = (double)(p.p1v * p.p3v), P2 = (double)(p.p2v * p.p3v);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        B[0][c] = PM[c] * P3 - PM[8 + c] * P1;
        B[1][c] = PM[4 + c] * P3 - PM[8 + c] * P2;
    }
    const double k2 = (double)d0, k4 = (double)d1, fx = (double)f;
    const double xy2 = x * x + y * y, xy4 = xy2 * xy2;
    const double xxpc = k4 * xy4 + x * (2 * k2 * x + 4 * k4 * x * xy2) + k2 * xy2 + 1.;
    const double xypc = x * (4 * k4 * xy2 * y + 2 * k2 * y);
    const double yxpc = y * (4 * k4 * x * xy2 + 2 * k2 * x);
    const double yypc = y * (2 * k2 * y + 4 * k4 * xy2 * y) + k4 * xy4 + k2 * xy2 + 1.;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double xpo = B[0][c], ypo = B[1][c];
        double xp = xxpc * xpo + xypc * ypo;
        double yp = yypc * ypo + yxpc * xpo;
        xp = xp * fx + yp * 0.;
        yp *= fx;
        B[0][c] = xp;
        B[1][c] = yp;
    }
}

}  // namespace

// CalcLSAXMulti (rpbacore.cpp:536-773), one thread per point.  The
// per-entry projections, residuals and derivatives are recomputed from the
// float point of the last trial instead of kept in per-entry arrays: the same
// operations on the same values, so the same numbers.  The current weights
// (reduced by robust flags) and the deletion flags live in per-entry
// workspace at the entry's track position.
__global__ __launch_bounds__(128) void k_lsax(LsaxArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.npts) return;
    const int n = a.len[i];
    if (n == 0) {
        a.nout[i] = 0;
        a.sg[i] = 0.;
        a.rd[i] = 0;
        return;
    }
    const long long b = a.base[i];
    int *fo = a.trk + b;
    float *wk = a.wk + 3 * b;
    unsigned char *dl = a.dl + b;
    const long long s0 = a.soff[i];
    const int ns = (int)(a.soff[i + 1] - s0);
    double *nv = a.nv + 9 * s0;
    int iter = 1, iflag = a.robustmode > 0 ? 2 : 1;
    float sig0 = 1.e20f, sig0opt;
    float X0 = (float)a.X[3 * (size_t)i], X1 = (float)a.X[3 * (size_t)i + 1], X2 = (float)a.X[3 * (size_t)i + 2];
    float Xold0 = X0, Xold1 = X1, Xold2 = X2;
    const float X0in = X0, X1in = X1, X2in = X2;
    float Xt0 = X0, Xt1 = X1, Xt2 = X2;  // the point of the last cost evaluation
    int convflag = 0, pxsd = n, pxs2 = 2 * n;
    double sgi = 0.;
    int rdi = 0;
    float wsum = 0.f;
    for (int js = 0; js < n; ++js) {
        const int q = fo[js];
        const float w0 = a.ow ? a.ow[3 * (size_t)q] : 1.f, w1 = a.ow ? a.ow[3 * (size_t)q + 1] : 1.f;
        const float w2 = a.ow ? a.ow[3 * (size_t)q + 2] : 0.f;
        wk[3 * js] = w0;
        wk[3 * js + 1] = w1;
        wk[3 * js + 2] = w2;
        dl[js] = 0;
        wsum += w0 + w1;
    }
    // CalcCorrPointMulti: float cost under the current weights
    auto cost = [&](double x0, double x1, double x2) {
        float s = 0.f;
        for (int js = 0; js < n; ++js) {
            const int q = fo[js], j = a.obs_cam[q];
            const LProj p = lsax_proj(a.PMd + 12 * (size_t)j, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], a.oxy[2 * (size_t)q],
                                      a.oxy[2 * (size_t)q + 1], x0, x1, x2);
            const float si = (float)(p.dx * p.dx * wk[3 * js] + p.dy * p.dy * wk[3 * js + 1] + 2 * p.dx * p.dy * wk[3 * js + 2]);
            s += si;
        }
        return s;
    };
    // the robust key of an entry (original weights) at the last trial
    auto key = [&](int js) {
        const int q = fo[js], j = a.obs_cam[q];
        const LProj p = lsax_proj(a.PMd + 12 * (size_t)j, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], a.oxy[2 * (size_t)q],
                                  a.oxy[2 * (size_t)q + 1], Xt0, Xt1, Xt2);
        const float w0 = a.ow ? a.ow[3 * (size_t)q] : 1.f, w1 = a.ow ? a.ow[3 * (size_t)q + 1] : 1.f;
        const float w2 = a.ow ? a.ow[3 * (size_t)q + 2] : 0.f;
        return (float)(p.dx * p.dx * w0 + p.dy * p.dy * w1 + 2 * p.dx * p.dy * w2);
    };
    sig0opt = cost(X0, X1, X2);
    if (a.weightonly) {
        wsum /= (float)pxs2;
        sgi += sig0opt / wsum;
        rdi += pxs2 - 3;
    }
    while (iflag > -1 && iter < 101) {
        double N[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
        if (a.wxyflag)
            for (int e = 0; e < 9 * ns; ++e) nv[e] = 0.;
        for (int js = 0; js < n; ++js) {
            const int q = fo[js], j = a.obs_cam[q];
            const double *PM = a.PMd + 12 * (size_t)j;
            const LProj p = lsax_proj(PM, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], a.oxy[2 * (size_t)q], a.oxy[2 * (size_t)q + 1],
                                      Xt0, Xt1, Xt2);
            double B[2][3];
            lsax_jac(PM, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], p, B);
            const float wx = wk[3 * js], wy = wk[3 * js + 1], wxy = wk[3 * js + 2];
            double bw[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                bw[c] = B[0][c] * wx + B[1][c] * wxy;
                bw[3 + c] = B[1][c] * wy + B[0][c] * wxy;
            }
            double T[9];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c) T[r * 3 + c] = B[0][r] * bw[c] + B[1][r] * bw[3 + c];
#pragma unroll
            for (int e = 0; e < 9; ++e) N[e] += T[e];
            if (a.wxyflag && a.P > 1) {
                const int ij = a.part[j];
                for (int s = 0; s < ns; ++s)
                    if (a.pit[s0 + s] && a.spart[s0 + s] != ij)
                        for (int e = 0; e < 9; ++e) nv[9 * s + e] += T[e];
            }
        }
        const double m1 = 1.001;
        N[0] *= m1;
        N[4] *= m1;
        N[8] *= m1;
        double Ni[9];
        pinv3(N, Ni);
        // exip += (eNppi * Bw^T) * corr, left to right (rpbacore.cpp:673-678)
        double ex[3] = {0., 0., 0.};
        for (int js = 0; js < n; ++js) {
            const int q = fo[js], j = a.obs_cam[q];
            const double *PM = a.PMd + 12 * (size_t)j;
            const LProj p = lsax_proj(PM, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], a.oxy[2 * (size_t)q], a.oxy[2 * (size_t)q + 1],
                                      Xt0, Xt1, Xt2);
            double B[2][3];
            lsax_jac(PM, a.cf[j], a.cd[2 * j], a.cd[2 * j + 1], p, B);
            const float wx = wk[3 * js], wy = wk[3 * js + 1], wxy = wk[3 * js + 2];
            double bw[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                bw[c] = B[0][c] * wx + B[1][c] * wxy;
                bw[3 + c] = B[1][c] * wy + B[0][c] * wxy;
            }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const double M0 = Ni[r * 3 + 0] * bw[0] + Ni[r * 3 + 1] * bw[1] + Ni[r * 3 + 2] * bw[2];
                const double M1 = Ni[r * 3 + 0] * bw[3] + Ni[r * 3 + 1] * bw[4] + Ni[r * 3 + 2] * bw[5];
                ex[r] += M0 * p.dx + M1 * p.dy;
            }
        }
        X0 = Xold0 + (float)ex[0];
        X1 = Xold1 + (float)ex[1];
        X2 = Xold2 + (float)ex[2];
        sig0 = cost(X0, X1, X2);
        Xt0 = X0;
        Xt1 = X1;
        Xt2 = X2;
        if (iflag == 1 && a.robustmode == 2) {
            for (int js = 0; js < n; ++js) {
                const int j = a.obs_cam[fo[js]];
                if (key(js) * a.sig0wv[j] > 16.f) {
                    wk[3 * js] *= 1.e-4f;
                    wk[3 * js + 1] *= 1.e-4f;
                    wk[3 * js + 2] *= 1.e-4f;
                    if (wk[3 * js] + wk[3 * js + 1] < 2.e-8f) {
                        wk[3 * js] = wk[3 * js + 1] = 1.e-8f;
                        wk[3 * js + 2] = 0.f;
                    }
                    --pxsd;
                    dl[js] = 1;
                }
            }
            if (pxsd < 2) break;
        }
        if ((double)sig0 < (double)sig0opt * 1.001) {
            if (sig0opt < sig0 * 1.05f) --iflag;
            if (!a.weightonly) {
                a.X[3 * (size_t)i] = Xold0 = X0;
                a.X[3 * (size_t)i + 1] = Xold1 = X1;
                a.X[3 * (size_t)i + 2] = Xold2 = X2;
            }
            sig0opt = sig0;
            convflag = 1;
            if (a.wxyflag)
                for (int s = 0; s < ns; ++s) {
                    const double *E = nv + 9 * s;
                    float *w = a.w3 + 6 * (s0 + s);
                    w[0] = (float)E[0];
                    w[1] = (float)E[4];
                    w[2] = (float)E[8];
                    w[3] = (float)E[1];
                    w[4] = (float)E[2];
                    w[5] = (float)E[5];
                }
        } else {
            --iflag;
        }
        if (a.weightonly) break;
        ++iter;
    }
    if (a.robustmode == 1 && a.keys)
        for (int js = 0; js < n; ++js) a.keys[b + js] = key(js);
    int nk = n;
    if (a.robustmode == 2) {
        int w = 0, ndel = 0;
        for (int js = 0; js < n; ++js) {
            if (dl[js]) {
                ++ndel;
                continue;
            }
            fo[w++] = fo[js];
        }
        nk = w;
        pxs2 -= 2 * ndel;
    }
    if (nk < 2 || !convflag) {
        nk = 0;
        a.X[3 * (size_t)i] = X0in;
        a.X[3 * (size_t)i + 1] = X1in;
        a.X[3 * (size_t)i + 2] = X2in;
    } else if (!a.weightonly) {
        wsum /= (float)pxs2;
        sgi += sig0opt / wsum;
        rdi += pxs2 - 3;
    }
    a.nout[i] = nk;
    a.sg[i] = sgi;
    a.rd[i] = rdi;
}

void launch_lsax(const LsaxArgs &a, hipStream_t st)
{
    if (a.npts <= 0) return;
    k_lsax<<<(unsigned)((a.npts + 127) / 128), 128, 0, st>>>(a);
}

// ---------------------------------------------------------------------------
// block-segmented PCG
// ---------------------------------------------------------------------------

// the done counter, the all-done flag of the union SpMV and the strips' live
// counts cleared (a kernel, so the set-up can be recorded into a graph)
__global__ void k_seg_reset(PcgArgs a, SegArgs s)
{
    if (threadIdx.x == 0) {
        *s.ndone = 0;
        a.S[0].done = 0;
    }
    if (s.strip_live)
        for (int i = threadIdx.x; i < s.nstrip; i += blockDim.x) s.strip_live[i] = 0;
}

// the strips holding cameras c0 .. c1 - 1 of a block: add v to their live
// counts (threads of the workgroup share the strips)
__device__ __forceinline__ void seg_strips_add(const SegArgs &s, int c0, int c1, int v)
{
    if (!s.strip_live || c1 <= c0) return;
    for (int st = c0 / kSlab + (int)threadIdx.x; st <= (c1 - 1) / kSlab; st += blockDim.x)
        atomicAdd(s.strip_live + st, v);
}

// per block: x = 0, r = b, best = b, h = C r, d = h, rho = r.h (rbacore.cpp:1829-1847).
// A block whose msize the host set to 0 does not run: it enters done with a
// zero step (x = best = d = 0) and counts as stopped.  The block that makes the
// count reach nb raises a.S[0].done, which ends the union SpMV launches that
// follow.
__global__ __launch_bounds__(256) void k_seg_init(PcgArgs a, SegArgs s)
{
    __shared__ double sh[4];
    const int bb = blockIdx.x;
    const int c0 = s.boff[bb], c1 = s.boff[bb + 1];
    const size_t j0 = 9 * (size_t)c0, j1 = 9 * (size_t)c1;
    if (s.S[bb].msize <= 0) {
        for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
            a.x[j] = 0.;
            a.best[j] = 0.;
            a.d[j] = 0.;
        }
        if (threadIdx.x == 0) {
            SegScal z = s.S[bb];
            z.rho = 0.;
            z.m0 = z.mold = 1.e50;
            z.k = 0;
            z.done = 1;
            z.iters = 0;
            s.S[bb] = z;
            if (atomicAdd(s.ndone, 1) + 1 == s.nb) a.S[0].done = 1;
        }
        return;
    }
    seg_strips_add(s, c0, c1, 1);
    for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        a.x[j] = 0.;
        a.r[j] = a.b[j];
        a.best[j] = a.b[j];
    }
    __syncthreads();
    double rh = 0.;
    for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const size_t k = j / 9, row = j % 9;
        const double *C = a.Cinv + 81 * k + 9 * row;
        const double *rr = a.r + 9 * k;
        double h = 0.;
#pragma unroll
        for (int c = 0; c < 9; ++c) h += C[c] * rr[c];
        a.h[j] = h;
        a.d[j] = h;
        rh += a.r[j] * h;
    }
    rh = pblock_sum(rh, sh);
    if (threadIdx.x == 0) {
        SegScal z = s.S[bb];  // msize set by the host
        z.rho = rh;
        z.m0 = 1.e50;
        z.mold = 1.e50;
        z.k = 0;
        z.done = 0;
        z.iters = 0;
        s.S[bb] = z;
    }
}

// (S d)_k for every camera of a running block: the diagonal product and the
// row partials of its strip's chunks, then the column shares of its slabs,
// in the order of the global solve's gather; d_k.(S d)_k
__global__ __launch_bounds__(256) void k_seg_gather(PcgArgs a, SegArgs s)
{
    __shared__ double red[28][9];
    const int k = blockIdx.x;
    if (s.S[s.cblk[k]].done) return;
    const int t = threadIdx.x;
    const int g = t / 9, r = t % 9;
    if (g < 28) {
        const int st = k / kSlab, gk = k - kSlab * st;
        const int c0 = a.chunk_off[st], nck = a.chunk_off[st + 1] - c0;
        double acc = (g == 0) ? a.Ad[9 * (size_t)k + r] : 0.;
        for (int it = g; it < nck; it += 28) acc += a.rowpart[63 * (size_t)(c0 + it) + 9 * gk + r];
        const int q0 = a.cs_off[k], nq = a.cs_off[k + 1] - q0;
        const int *cl = a.cs_list + q0;
        for (int it = g; it < nq; it += 28) acc += a.colpart[9 * (size_t)cl[it] + r];
        red[g][r] = acc;
    }
    __syncthreads();
    if (t < 9) {
        double v = 0.;
        for (int gg = 0; gg < 28; ++gg) v += red[gg][t];
        s.Sdf[9 * (size_t)k + t] = v;
        red[0][t] = v * a.d[9 * (size_t)k + t];
    }
    __syncthreads();
    if (t == 0) {
        double e = 0.;
        for (int q = 0; q < 9; ++q) e += red[0][q];
        s.e[k] = e;
    }
}

// one PCG iteration of a block (rbacore.cpp:1851-1911): alpha = rho / d.S d;
// x += alpha d; r -= alpha S d and m = |alpha S d|^2; h1 = C r; the best
// iterate; the stopping rule m 1e5 < m0; beta; d = h1 + beta d
__global__ __launch_bounds__(1024) void k_seg_step(PcgArgs a, SegArgs s)
{
    __shared__ double sh[16];
    const int bb = blockIdx.x;
    const SegScal S = s.S[bb];
    if (S.done) return;
    const int c0 = s.boff[bb], c1 = s.boff[bb + 1];
    const size_t j0 = 9 * (size_t)c0, j1 = 9 * (size_t)c1;
    double dAd = 0.;
    for (int k = c0 + threadIdx.x; k < c1; k += blockDim.x) dAd += s.e[k];
    dAd = pblock_sum(dAd, sh);
    const double rho = S.rho;
    const double alpha = (fabs(rho) < 1.e-20 || fabs(rho) > 1.e20) ? 0. : rho / dAd;
    double m = 0.;
    for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        a.x[j] += alpha * a.d[j];
        const double dr = alpha * s.Sdf[j];
        a.r[j] = a.r[j] - dr;
        m += dr * dr;
    }
    __syncthreads();
    double num = 0., rh = 0.;
    for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
        const size_t k = j / 9, row = j % 9;
        const double *C = a.Cinv + 81 * k + 9 * row;
        const double *rr = a.r + 9 * k;
        double h1 = 0.;
#pragma unroll
        for (int c = 0; c < 9; ++c) h1 += C[c] * rr[c];
        a.h1[j] = h1;
        num += a.r[j] * (h1 - a.h[j]);
        rh += a.r[j] * h1;
    }
    m = pblock_sum(m, sh);
    num = pblock_sum(num, sh);
    rh = pblock_sum(rh, sh);
    const int it = S.k;
    const double m0 = (it == 0) ? m : S.m0;
    const bool better = m < S.mold;
    if (better)
        for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) a.best[j] = a.x[j];
    SegScal z = S;
    z.m0 = m0;
    if (better) z.mold = m;
    if (m * 1.e5 < m0) {
        z.done = 1;
        z.iters = it + 1;
    } else {
        double beta = (fabs(num) < 1.e-10) ? 0. : num / rho;
        if (beta < 0.) beta = 0.;
        for (size_t j = j0 + threadIdx.x; j < j1; j += blockDim.x) {
            const double h1 = a.h1[j];
            a.d[j] = h1 + beta * a.d[j];
            a.h[j] = h1;
        }
        z.rho = rh;
        z.k = it + 1;
        if (it + 1 >= S.msize) {
            z.done = 1;
            z.iters = S.msize;
        }
    }
    __syncthreads();
    if (z.done) seg_strips_add(s, c0, c1, -1);
    if (threadIdx.x == 0) {
        s.S[bb] = z;
        if (z.done && atomicAdd(s.ndone, 1) + 1 == s.nb) a.S[0].done = 1;
    }
}

void launch_seg_init(const PcgArgs &a, const SegArgs &s, hipStream_t st)
{
    if (s.nb <= 0) return;
    k_seg_reset<<<1, 256, 0, st>>>(a, s);
    k_seg_init<<<s.nb, 256, 0, st>>>(a, s);
}

void launch_seg_update(const PcgArgs &a, const SegArgs &s, hipStream_t st)
{
    if (s.nb <= 0) return;
    k_seg_gather<<<s.nc, 256, 0, st>>>(a, s);
    k_seg_step<<<s.nb, 1024, 0, st>>>(a, s);
}

void launch_seg_iter(const PcgArgs &a, const SegArgs &s, hipStream_t st)
{
    if (s.nb <= 0) return;
    launch_spmv(a, st);
    launch_seg_update(a, s, st);
}

__global__ void k_seg_nan(int nc, int npts, const int *cblk, const int *ptblk, const double *sol, const double *Xn,
                          int *flag)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nc) {
        bool bad = false;
        for (int q = 0; q < 9; ++q) {
            const float xf = (float)sol[9 * (size_t)t + q];  // the camera step as the reference casts it (rbacore.cpp:1185)
            if (isnan(xf) || isinf(xf)) bad = true;
        }
        if (bad) flag[cblk[t]] = 1;
    } else if (t < nc + npts) {
        const int i = t - nc;
        bool bad = false;
        for (int q = 0; q < 3; ++q) {
            const double v = Xn[3 * (size_t)i + q];
            if (isnan(v) || isinf(v)) bad = true;
        }
        if (bad) flag[ptblk[i]] = 1;
    }
}

void launch_seg_nan(int nc, int npts, const int *cblk, const int *ptblk, const double *sol, const double *Xn, int nb,
                    int *flag, hipStream_t st)
{
    (void)hipMemsetAsync(flag, 0, sizeof(int) * (size_t)(nb > 0 ? nb : 1), st);
    const long long n = (long long)nc + npts;
    if (n <= 0) return;
    k_seg_nan<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(nc, npts, cblk, ptblk, sol, Xn, flag);
}

}  // namespace rpba
